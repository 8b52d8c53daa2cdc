"""MI355X line-intensity engine (gfx950): host-side Python helpers.

The hot path lives in the C-ABI library built from csrc/ (libgi_hip.so);
binding.py loads it with ctypes. synth.py builds seeded synthetic inputs.
The directory name has a hyphen, so it is loaded by path under the module
name ``gpu_intensities_amd`` (see load_package in __graft_entry__.py).
"""
