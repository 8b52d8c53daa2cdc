/*
 * gpu_intensities.h — C ABI of the MI355X (gfx950) line-intensity engine.
 *
 * Plain C: POD structs, pointers and sizes only. Every entry point returns a
 * gi_status; gi_last_error() gives the message. Nothing calls exit().
 *
 * The reference (ahmed-f-alrefaie/gpu-intensities) has no FFI layer: its
 * boundary is the host-callable CUDA API of cuda_host.cuh:10-24 and
 * dipole_kernals.cuh:4-13, driven by dipole_do_intensities_async_omp
 * (cuda_host.cu:995-1506). Each declaration below names the reference
 * interface it replaces.
 *
 * Threading: a context is used by one host thread at a time. All device
 * memory is owned by the context and freed by gi_destroy().
 */
#ifndef GPU_INTENSITIES_H
#define GPU_INTENSITIES_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gi_ctx gi_ctx;

typedef enum gi_status {
    GI_OK = 0,
    GI_EINVAL = 1,     /* bad argument or inconsistent input                 */
    GI_EHIP = 2,       /* a HIP runtime call failed                          */
    GI_ENOMEM = 3,     /* device or host allocation failed                   */
    GI_ESTATE = 4,     /* call out of order (e.g. compute before prepare)    */
    GI_ECAPACITY = 5,  /* output arrays too small; required size is returned */
    GI_EBUDGET = 6,    /* request does not fit the context's memory budget   */
    GI_ETIMEOUT = 7    /* gi_serve_wait: the request has not finished yet     */
} gi_status;

/* Arithmetic of the half-linestrength and correlation products. Energies,
 * Boltzmann factors and the intensity step are FP64 in every mode. */
typedef enum gi_mode {
    GI_MODE_FP64 = 0,          /* plain FP64: one GEMV-form product per lower level */
    GI_MODE_FP64_BATCHED = 1,  /* FP64 matrix-core products over batches of lower levels */
    GI_MODE_SPLIT = 2          /* batched, both products as split int8 digits on the
                                  matrix cores (see gi_set_split_format)          */
} gi_mode;

/* Basis of one J: the TO_bset_contrT fields (trove_objects.h:49-72) read by
 * the path, after correlate_index (trove_functions.cpp:680-858) and
 * compute_ijterms (trove_functions.cpp:927-966).
 * Replaces create_and_copy_bset_contr_to_gpu (cuda_host.cu:368-376) and the
 * device struct cuda_bset_contrT (cuda_objects.cuh:5-21). */
typedef struct gi_bset {
    int jval;
    int Maxsymcoeffs;             /* cases                                   */
    int Maxcontracts;             /* dimen: primitive functions              */
    int mat_size;                 /* third extent of repres                  */
    const int *icontr2icase;      /* [2*dimen]: icase at r, ilambda at r+dimen */
    const int *iroot_correlat_j0; /* [dimen]: J=0 (vibrational) index        */
    const int *k;                 /* [dimen]                                 */
    const int *ktau;              /* [dimen]: 2k + tau                       */
    const int *N;                 /* [nrep*Maxsymcoeffs]: irr[g].N[c] at g*Maxsymcoeffs + c */
    const int *Ntotal;            /* [nrep]                                  */
    const int *ijterms;           /* [nrep*Maxsymcoeffs]: c + g*Maxsymcoeffs */
    const double *const *repres;  /* [nrep] -> Ntotal[g]*degen[g]*mat_size   */
    const int *nsize;             /* [nrep]: eigenvector length per symmetry */
} gi_bset;

/* Job parameters: FintensityJob (fields.h:48-82) as read by the deck parser
 * (fields.cpp:145-282). quanta_* hold lo at [i] and hi at [nmodes+i]. */
typedef struct gi_params {
    int nrepres;
    const int *degen;             /* [nrepres]                               */
    int sym_type;                 /* 1 C2V(M), 2 C3V(M), 3 D3H(M) (fields.h:9-11) */
    int nmodes;
    double zpe, temperature, q_stat;
    double freq_window[2];
    double erange_lower[2];
    double erange_upper[2];
    int jvals[2];
    const double *gns;            /* [nrepres]                               */
    const int *isym_pairs;        /* [nrepres] (SELECTION)                   */
    int reduced;                  /* SYMMETRY REDUCED                        */
    const int *quanta_lower;      /* [2*nmodes]                              */
    const int *quanta_upper;      /* [2*nmodes]                              */
    double thresh_linestrength;   /* emit only S >= this (0: all)            */
    double thresh_intensity;      /* emit only I >= this (0: all)            */
} gi_params;

/* Emitted transitions, structure of arrays with room for `capacity` rows.
 * amp holds W = m*m doubles per row, m the largest degeneracy of the
 * context's symmetry (gi_params.degen; the reference's sym_maxdegen): entry
 * idegI + m*idegF is the amplitude A(idegI, idegF), zero for components a
 * level does not have. This is the layout of the reference's line_str per
 * final state (cuda_host.cu:1467-1471: idegI + idegF*sym_maxdegen), so a
 * symmetry without degenerate irreps (C2v) carries one amplitude per line.
 * gi_amp_width gives W; amp needs room for capacity*W doubles. */
typedef struct gi_lines {
    long long capacity;
    long long n;                  /* out: rows written (or required)         */
    int *ilevelI, *ilevelF;       /* indices into the level table            */
    double *nu, *S, *A_einst, *I;
    double *amp;
} gi_lines;

/* Context on one device. Replaces cudaSetDevice + dipole_initialise_gpu's
 * device set-up (cuda_host.cu:995-1008, 900-909). */
int gi_create(int device, gi_ctx **out);
/* Frees every device allocation (the reference leaks them, cuda_host.cu:1501-1503). */
void gi_destroy(gi_ctx *ctx);
const char *gi_last_error(const gi_ctx *ctx);

/* Replaces read_fields' job fields + copy_intensity_info
 * (fields.cpp:236-282; dipole_kernals.cu:8-14) and find_igamma_pair
 * (trove_functions.cpp:1041-1072; an ambiguous pairing is GI_EINVAL). */
int gi_set_params(gi_ctx *ctx, const gi_params *p);

/* J list of the context. gi_set_params sets the deck's two J values
 * (jvals[0], jvals[1]); this replaces them with nj ascending values, so that
 * one context holds the bases, eigenvector blocks and levels of many J and
 * computes every pair with |J' - J| <= 1 in the same launches. The reference
 * holds two J per run (fields.h:61 bset_contr[3], nJ = 2 at
 * cuda_host.cu:1005) and is run once per (J, J+1); with the Q-branch rule
 * (trove_functions.cpp:1457: no J = J' = jvals[0]) the list jvals[0..nj-1]
 * gives exactly the union of the runs (jvals[i], jvals[i+1]), i < nj - 1.
 * Call after gi_set_params; drops the bases, eigenvector blocks and level
 * table (their jind index this list). nj: 1..128. */
int gi_set_jvals(gi_ctx *ctx, int nj, const int *jvals);

/* Basis of jvals[jind] (jind = 0, 1, or an index into the gi_set_jvals
 * list). Replaces create_and_copy_bset_contr_to_gpu (cuda_host.cu:368-376). */
int gi_upload_bset(gi_ctx *ctx, int jind, const gi_bset *b);

/* Flipped dipole mu_a(vF, vI) at vF + vI*nvib + a*nvib^2, a = x, y, z.
 * Replaces read_dipole_flipped + copy_array_to_gpu(dipole_me)
 * (trove_functions.cpp:113-205; cuda_host.cu:983). on_device != 0 means
 * `dipole` is a device pointer on the context's device. */
int gi_upload_dipole(gi_ctx *ctx, int nvib, const double *dipole, int on_device);

/* Device bytes allowed for mu (0, the default: no limit), applied by the next
 * gi_upload_dipole. A larger mu stays in host memory: gi_prepare keeps its
 * permuted copy in pinned host memory, and every mode copies it to the device
 * in parts of floor(bytes / (24 * nvib)) vF rows (two parts on the device at
 * a time, the next copied while the current one computes), computing the
 * half linestrength rows of each part from it (the same sums in the same
 * order as from resident mu, so the same bits): GI_MODE_FP64 once per batch
 * of up to 256 lower levels, the matrix-core modes once per pipeline round.
 * GI_MODE_SPLIT builds mu's digit planes part by part at gi_prepare_split and
 * keeps them in pinned host memory, streaming their parts the same way.
 * Replaces the reference's split dipole (max_dipole_size, cuda_host.cu:10;
 * read_dipole_flipped_blocks, trove_functions.cpp:207-318;
 * do_1st_half_ls_blocks, cuda_host.cu:783-817), which re-copies every part
 * for every lower level. Sweeps and gi_half_ls need mu resident (GI_ESTATE).
 * GI_EINVAL from gi_upload_dipole when the budget is below one row
 * (24 * nvib bytes). */
int gi_set_mu_budget(gi_ctx *ctx, long long bytes);
/* Parts mu is streamed in (1: resident; 0: no mu uploaded). */
int gi_mu_parts(const gi_ctx *ctx);

/* Device bytes allowed for the eigenvector blocks uploaded from host memory
 * (0, the default: no limit), applied by the next gi_upload_vectors calls: a
 * block that would take the total over it stays in pinned host memory, and
 * every mode (also for requests) copies the records each chunk of lower
 * levels needs (its lower records and the upper records of its frequency
 * window; in GI_MODE_SPLIT the upper records' digit rows, which
 * gi_prepare_split builds on the device and keeps in pinned host memory for
 * such a block) to the device while the previous chunk computes. The lines
 * are the same bits as with resident blocks. Replaces the reference's
 * upper-vector batches sized to free device memory and read from disk for
 * every lower level (cuda_host.cu:1011-1027, 1235-1372). Captured sweeps need
 * every block resident (GI_ESTATE). Blocks passed as device pointers are
 * always resident. */
int gi_set_vector_budget(gi_ctx *ctx, long long bytes);
/* Number of eigenvector blocks kept in host memory. */
int gi_streamed_blocks(const gi_ctx *ctx);

/* Level table, energy sorted, structure of arrays (TO_PTeigen,
 * trove_objects.h:74-88, as filled by read_eigenvalues,
 * trove_functions.cpp:1074-1429). irec0 is the 0-based eigenvector record.
 * quanta is [n*(nmodes+1)]. */
int gi_set_levels(gi_ctx *ctx, int n, const double *energy, const int *jind,
                  const int *igamma, const int *ndeg, const int *irec0, const int *quanta);

/* Eigenvector block of (jvals[jind], igamma): nrec records of nsize doubles,
 * record irec0 at irec0*nsize. Kept resident in HBM. Replaces the per-level
 * fseek/fread + cublasSetVector / cudaMemcpyAsync of cuda_host.cu:1196-1199,
 * 1262-1273. on_device != 0: `vectors` is a device pointer that the context
 * adopts by reference (the caller keeps it alive and unchanged). */
int gi_upload_vectors(gi_ctx *ctx, int jind, int igamma, int nsize, int nrec,
                      const double *vectors, int on_device);

/* Builds the device plan (3j table, slot grids, symmetrisation maps,
 * coefficient tables, per-block energy order). Call after the uploads. */
int gi_prepare(gi_ctx *ctx);

/* The hot path for the given lower levels (indices into the level table;
 * NULL = all): half linestrengths, correlation with every upper level that
 * passes the window and selection filters, degeneracy sum, Einstein A and
 * intensity, thresholds, compaction and copy to `out` (host memory).
 * Replaces the body of the lower-level loop, cuda_host.cu:1174-1488
 * (do_1st_half_ls :767-780, correlate + cublasDdot :1284-1322, intensity
 * :1376-1407). Row order: by lower block (J, symmetry) and upper J; within
 * that, by lower level (in lower_ids order for GI_MODE_FP64, in energy order
 * for GI_MODE_FP64_BATCHED); within a lower level, by upper level energy.
 * If out->capacity is too small, returns GI_ECAPACITY with the required
 * count in out->n and writes nothing. With capacity for every candidate
 * (gi_count_candidates), the batched mode copies lines to `out` while it
 * still computes; `out` should then be pinned host memory. */
int gi_compute(gi_ctx *ctx, int mode, int nlower, const int *lower_ids, gi_lines *out);

/* Upper bound of the rows gi_compute can emit for these lower levels. */
int gi_count_candidates(gi_ctx *ctx, int nlower, const int *lower_ids, long long *count);

/* Doubles per row of gi_lines.amp for this context (1 or 4; see gi_lines),
 * fixed by gi_set_params; 0 for a null context or before gi_set_params. Replaces the reference's sym_maxdegen**2 stride of
 * line_str (cuda_host.cu:1467-1471). */
int gi_amp_width(const gi_ctx *ctx);

/* do_1st_half_ls (cuda_host.cu:767-780): correlate the symmetrised lower
 * vector vec (nsize of (jvals[jindI], igammaI), host memory) for component
 * ideg, then the half linestrength towards jvals[jindF]; h (host memory)
 * receives dimen(jvals[jindF]) values in the primitive order of the basis. */
int gi_half_ls(gi_ctx *ctx, int jindI, int jindF, int igammaI, int ideg,
               const double *vec, double *h);

/* gi_half_ls without allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same arguments,
 * result and errors as gi_half_ls, but every device buffer (the lower vector,
 * G, H and the result) lives in one static 64 MiB device array per GPU, so the
 * call makes no hipMalloc, hipFree or other allocation. GI_EINVAL when the
 * job needs more than the array holds (about Nvib <= 25,000). Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (several host
 * threads, or contexts used concurrently) race. */
int gi_half_ls_noalloc(gi_ctx *ctx, int jindI, int jindF, int igammaI, int ideg,
                       const double *vec, double *h);

/* device_correlate_vectors (dipole_kernals.cu:38-71): out (host, dimen of
 * jvals[jind]) = symmetrised -> primitive transform of vec (host, nsize). */
int gi_correlate(gi_ctx *ctx, int jind, int igamma, int ideg, const double *vec, double *out);

/* precompute_threej (trove_functions.cpp:899-925): out has (jmax+1)^2*9
 * doubles; entries outside the valid (k, k') range are zero. */
int gi_threej_table(int jmax, double *out);

/* Per-kernel timing of the last gi_compute, from HIP events recorded on the
 * stream the kernels run on. names/avg_ms/launches receive up to `max`
 * entries; *n is the number of kernels. */
int gi_kernel_stats(gi_ctx *ctx, int max, const char **names, double *total_ms,
                    int *launches, int *n);

/* Work counters of the last gi_compute: [0] half-ls units (lower level x J'
 * x component), [1] candidates, [2] emitted lines, [3] dense half-ls flops
 * (2*C*Nvib^2 per unit, C = component-minimum count), [4] half-ls flops
 * executed by the symmetry-blocked products, [5] correlation bytes
 * (8*nsize_F per candidate), [6] mu bytes read by the half-ls products
 * (per unit in GI_MODE_FP64, per batch in GI_MODE_FP64_BATCHED), [7] dense mu
 * bytes (24*Nvib^2 per unit), [8] correlation flops executed; host side of
 * the call in ms: [9] planning (items, chunks, metadata), [10] enqueueing the
 * launches, [11] waiting for the device and the copies after that, [13] the
 * whole call; [12] bytes of the lines copied to `out`; [14] tiles of the
 * half-ls products and [15] of the correlation products (matrix-core modes,
 * summed over launches); [16] useful correlation flops (2*nsize_F per
 * computed component pair of a candidate). n: up to 17. */
int gi_work_stats(gi_ctx *ctx, double *out, int n);

/* GI_MODE_SPLIT (no reference counterpart): every operand row or column of
 * the two products is scaled by a power of two and cut into `digits` signed
 * int8 digits of radix R; the digit-pair products (digits*(digits+1)/2 per
 * multiply-add) are summed exactly in int32 on the int8 matrix cores and
 * combined in FP64, so the relative error of a product is about R^-digits of
 * its operands' row/column maxima. R = 254 (digits in [-127, 127], default 5
 * digits: ~2^-40) or R = 128 (digits in [-64, 64], default 6: ~2^-42); a sum is
 * exact while digits * (R/2)^2 * K < 2^31 for the product length K (a class of
 * the half linestrength, a split-K part of the correlation), and the format is
 * chosen at gi_prepare_split: R = 254 when both products allow it, else 128.
 * Either default keeps S within 1e-6 relative of FP64 for lines down to 1e-6 of
 * the strongest. Energies, Boltzmann factors, Q, the fold and the intensity
 * step stay FP64. gi_prepare_split builds the resident digit planes of mu and
 * of every eigenvector block (digits bytes per element); gi_compute does it on
 * first use, the request server needs it beforehand. */
/* radix: 0 (automatic), 128 or 254; digits per product: 0 (the default of the
 * radix) or 3..6. GI_EINVAL from gi_prepare_split when no allowed format keeps
 * the sums exact. */
int gi_set_split_format(gi_ctx *ctx, int radix, int halfls_digits, int corr_digits);
/* The format of the resident digit planes (that of the last gi_prepare_split;
 * before one, radix 254 with 5 digits). */
int gi_get_split_format(gi_ctx *ctx, int *radix, int *halfls_digits, int *corr_digits);
/* gi_set_split_format with the radix request kept: `digits` (3..6) for both
 * products, or one count per product, the half linestrength (mu and G) and the
 * correlation (eigenvector blocks and folded vectors). */
int gi_set_split_digits(gi_ctx *ctx, int digits);
int gi_set_split_digits_each(gi_ctx *ctx, int halfls_digits, int corr_digits);
int gi_prepare_split(gi_ctx *ctx);

/* Tile shapes of the two matrix-core kernels of GI_MODE_FP64_BATCHED (no
 * reference counterpart: the reference has no batched form). cfg 0..3 =
 * rows x columns 64x64, 64x128, 128x64, 128x128; for the half linestrength
 * also 4 (the default): per pipeline round 128x128 on eight wavefronts
 * while the round has at least three waves of such tiles for the device,
 * else 64x64 (the last, small rounds), and 5: 128x128 on eight wavefronts instead of four (half
 * the accumulators per wavefront, twice the wavefronts per SIMD); -1 leaves
 * a kernel as it is. Results are bitwise the same
 * for every shape (the K order of each product does not depend on it). */
int gi_set_tiles(gi_ctx *ctx, int halfls_cfg, int corr_cfg);
int gi_get_tiles(gi_ctx *ctx, int *halfls_cfg, int *corr_cfg);

/* Launch configuration of the batched modes (no reference counterpart: the
 * reference has no batched form). halfls_cfg / corr_cfg are the tile shapes
 * of gi_set_tiles; pipe_rounds is the number of rounds a gi_compute with room
 * for every candidate is cut into, each round's lines being copied to the
 * host while later rounds compute (1..64; default 4): every round but the
 * last takes pipe_share percent of the work left (10..90; default 50), so the
 * copy after the last round is small. A share above 50 shrinks the last
 * round further and suits workloads whose line copy takes well under half of
 * the compute time. Lines are bitwise the same for every setting. */
typedef struct gi_tuning {
    int halfls_cfg;
    int corr_cfg;
    int pipe_rounds;
    int split_cfg;    /* GI_MODE_SPLIT half-ls tiles: 0 = 64 x 64, 1 = 128 x 64,
                         2 = 64 x 64 with two K steps of loads in flight,
                         3 = 128 x 64 with two K steps in flight (8 wavefronts),
                         4 = per round 3 when its tiles average at least 40 K
                         steps of 64, else 2 (the default) */
    int pipe_share;   /* percent of the remaining work per round but the last */
} gi_tuning;
int gi_set_tuning(gi_ctx *ctx, const gi_tuning *t);
int gi_get_tuning(gi_ctx *ctx, gi_tuning *t);

/* Routing thresholds of the batched modes (no reference counterpart), for
 * tests that reach at small sizes the kernels real sizes select: kpart_len,
 * the correlation K per split-K part in both matrix-core modes
 * (parts = min(16, K / kpart_len); default 1024 for FP64 and 2048 for the
 * split digits, at most 2048, at least 64); rec_lds, the doubles of an eigenvector record the
 * G builds stage in LDS (gathers beyond them read a longer record in global memory; default
 * 19968; below it the FP64 build also reads the full record map instead of
 * the compact one), or -1: no record staged, every chunk takes the
 * global-gather builds;
 * split2_steps / split2_segs, the largest class (64-byte K steps, segments)
 * the two-step split half-ls takes (larger ones run the one-step kernel;
 * default 512 / 64); fold_split_rows, the longest folded column the fused
 * fold-and-digit-split kernel takes (default 32768). 0 restores a default;
 * values above the default are GI_EINVAL (the kernels' own limits). The
 * partly staged and global-gather builds, the one-step split kernel and the two-kernel fold
 * give bitwise the same lines as the paths they replace; another split-K
 * part count changes the correlation's summation order. */
int gi_set_path_limits(gi_ctx *ctx, int kpart_len, int rec_lds, int split2_steps,
                       int split2_segs, int fold_split_rows);

/* Persistent rounds (no reference counterpart; the reference launches one
 * kernel per lower level and pair). on = 1: in GI_MODE_FP64_BATCHED every
 * pipeline round runs its half linestrength, fold, correlation and intensity
 * steps as one persistent launch after its G build: resident workgroups take
 * tasks in ticket order, stage by stage, and a task waits only for the
 * previous stage's tasks of its own chunk, so the stages of different chunks
 * overlap instead of each stage draining the device. on = 0 (the default):
 * one launch per stage. Lines are bitwise the same either way. Calls with mu
 * or eigenvector blocks in host memory use the per-stage launches. */
int gi_set_persist(gi_ctx *ctx, int on);
int gi_get_persist(gi_ctx *ctx, int *on);

/* Autotuner (no reference counterpart): runs gi_compute in `mode` over the
 * given lower levels (NULL = all) `reps` times per setting into `out`,
 * and keeps the fastest setting of each knob: the two tile shapes (FP64
 * batched mode) or the split half-ls tile shape (split mode), each judged by
 * its kernel's time, and the pipeline round count together with the share of
 * the work per round, as pairs (both matrix-core modes when `out` has room for
 * every candidate, judged by the call's wall time), then the FP64 half-ls tile once more (from
 * the setting in force at the call, so the first scan's pick holds no incumbency) at the
 * chosen rounds. Settings are scanned in a fixed order starting from
 * the current one, and a setting replaces the incumbent only when it is more
 * than 3% faster, so repeated runs pick the same table unless the timings
 * really differ. The chosen setting is applied and returned in *best;
 * `table` (cap bytes, NUL-terminated, *len = full length) receives one line
 * per measurement and a final "best ..." line, to be kept and re-applied with
 * gi_set_tuning. */
int gi_autotune(gi_ctx *ctx, int mode, int nlower, const int *lower_ids, gi_lines *out,
                int reps, gi_tuning *best, char *table, long long cap, long long *len);

/* Captured sweeps (no reference counterpart). gi_sweep_create plans one
 * gi_compute in a matrix-core mode over the given lower levels (NULL = all)
 * and records its launch sequence as one HIP graph per pipeline round,
 * without running it; the sweep owns a stream and scratch of its own.
 * gi_sweep_run replays the graphs with a new temperature, partition function
 * (<= 0: Q(T) from gi_partition_function) and thresholds: every product is
 * recomputed on the device, the host only launches the graphs and copies the
 * lines, so a series of runs (line lists at several temperatures, or the same
 * list again) skips the host planning and per-kernel launches of gi_compute.
 * `out` needs room for every candidate (gi_sweep_candidates); rows come in
 * the order of gi_compute. The other job parameters are those in force at
 * creation. While sweeps exist, calls that change the context (gi_set_params,
 * uploads, gi_prepare, gi_set_split_digits) return GI_ESTATE; gi_destroy
 * frees the context's sweeps, whose handles are then invalid. */
typedef struct gi_sweep gi_sweep;
int gi_sweep_create(gi_ctx *ctx, int mode, int nlower, const int *lower_ids, gi_sweep **out);
int gi_sweep_run(gi_sweep *sweep, double temperature, double q_stat,
                 double thresh_linestrength, double thresh_intensity, gi_lines *out);
long long gi_sweep_candidates(const gi_sweep *sweep);
int gi_sweep_destroy(gi_sweep *sweep);

/* Partition function over the resident level table:
 * Q(T) = sum_i gns(g_i) * ndeg_i * (2 J_i + 1) * exp(-(E_i - ZPE) * c2 / T),
 * c2 = planck * vellgt / boltz with the constants of cuda_host.cu:12-18.
 * The reference takes Q from the deck (fields.cpp:236-282, q_stat); this
 * gives the caller a value for other temperatures. Reduced on the device in
 * a fixed order (deterministic). */
int gi_partition_function(gi_ctx *ctx, double temperature, double *q);

/* ------------------------------------------------------------------------
 * Request server. The reference computes one deck job per process
 * (cuda_host.cu:995-1506); it has no request interface. These entry points
 * keep the prepared basis, dipole and eigenvectors resident and serve
 * requests against them: a worker thread admits queued requests while their
 * device memory fits the budget, batches the admitted ones (same mode) into
 * one launch sequence, and alternates two streams so that one batch's lines
 * are copied out while the next computes. A request's lines are the same
 * whatever it is batched with.
 * While the server runs, calls that change the context (gi_set_params,
 * uploads, gi_prepare) and gi_compute return GI_ESTATE. */
typedef struct gi_request {
    int jindI, igammaI, jindF;    /* lower block (J, symmetry) and upper J; upper
                                     symmetry = SELECTION partner of igammaI  */
    int mode;                     /* gi_mode                                  */
    double temperature;           /* K                                        */
    double q_stat;                /* partition function; <= 0: gi_partition_function(T) */
    double elow_lo, elow_hi;      /* lower levels with E - ZPE in [lo, hi], on top of
                                     the job's lower-level filters              */
    double nu_lo, nu_hi;          /* frequency window (replaces freq_window)  */
    double thresh_linestrength, thresh_intensity;
} gi_request;

/* Start the worker. budget_bytes bounds the device memory of the requests
 * in flight (candidate records and half-linestrength buffers). */
int gi_serve_start(gi_ctx *ctx, long long budget_bytes);
/* Queue a request; *ticket identifies it. GI_EBUDGET if it alone exceeds
 * the budget, GI_EINVAL if |J' - J| > 1 or a field is out of range. */
int gi_serve_submit(gi_ctx *ctx, const gi_request *req, long long *ticket);
/* Wait up to timeout_s (<= 0: no limit) for a request. Returns GI_OK with
 * the line count and submit-to-done latency, GI_ETIMEOUT, or the
 * request's own error. */
int gi_serve_wait(gi_ctx *ctx, long long ticket, double timeout_s, long long *nlines,
                  double *latency_ms);
/* Copy a finished request's lines to `out` (rows in the order of
 * gi_compute) and release the request. GI_ECAPACITY leaves it in place. */
int gi_serve_fetch(gi_ctx *ctx, long long ticket, gi_lines *out);
/* Batches launched, requests finished and the largest batch so far. */
int gi_serve_stats(gi_ctx *ctx, long long *batches, long long *served, long long *max_batch);
/* Finish every queued request, then stop the worker (also done by gi_destroy). */
int gi_serve_stop(gi_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
