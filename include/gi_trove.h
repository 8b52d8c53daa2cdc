/*
 * gi_trove.h — TROVE input files, the stdin deck and the line list, for the
 * MI355X line-intensity engine (gpu_intensities.h).
 *
 * The reference reads its job from stdin (read_fields, fields.cpp:145-282)
 * and TROVE checkpoints from the working directory (file names
 * trove_functions.cpp:24-28): j0_extfield.chk (dipole, :113-205),
 * j0eigen_quanta<J>.chk (basis, :401-679, then correlate_index :680-858 and
 * compute_ijterms :927-966), j0eigen_descr<J>_<G>.chk (levels,
 * :1074-1429) and j0eigen_vectors<J>_<G>.chk (eigenvectors, raw doubles
 * read per record at cuda_host.cu:1196). It prints one line per transition
 * (cuda_host.cu:1444-1473). These entry points do the same reading and
 * printing in front of and behind a gi_ctx; none of them calls exit().
 */
#ifndef GI_TROVE_H
#define GI_TROVE_H

#include "gpu_intensities.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gi_trove gi_trove;

/* Parse the deck text (the reference's stdin) and read every file it needs
 * from directory `dir`. Always returns a handle in *out (NULL only when out
 * of host memory); on failure the status is non-zero and gi_trove_error()
 * says why. Replaces dipole_initialise_cpu (cuda_host.cu:837-898). */
int gi_trove_open(const char *dir, const char *deck, gi_trove **out);
const char *gi_trove_error(const gi_trove *t);
void gi_trove_close(gi_trove *t);

/* gi_set_params, gi_upload_bset x2, gi_upload_dipole, gi_set_levels,
 * gi_upload_vectors for every read block, gi_prepare. The deck's THRESH_LINE
 * is not applied, as in the reference (parsed at fields.cpp:204-206, its cut
 * commented out at cuda_host.cu:1408), unless gi_trove_apply_thresh_line
 * asked for it: then lines with S below it are not emitted. */
int gi_trove_load(gi_trove *t, gi_ctx *ctx);
/* on != 0: gi_trove_load applies THRESH_LINE as a linestrength cut (no
 * reference counterpart; the reference never applies it). */
int gi_trove_apply_thresh_line(gi_trove *t, int on);

/* Views of what was read (valid until gi_trove_close). */
int gi_trove_params(const gi_trove *t, gi_params *p);
int gi_trove_bset(const gi_trove *t, int jind, gi_bset *b);
int gi_trove_dipole(const gi_trove *t, int *nvib, const double **flipped);
/* Level table in energy order: energy, jind, igamma, ndeg, irec0 (0-based
 * record), ilevel (0-based row of its descr file), quanta [n*(nmodes+1)],
 * krot, taurot. */
int gi_trove_levels(const gi_trove *t, int *n, const double **energy, const int **jind,
                    const int **igamma, const int **ndeg, const int **irec0,
                    const int **ilevel, const int **quanta, const int **krot,
                    const int **taurot);
/* Eigenvector block of (jind, igamma) as read: nrec records of nsize. */
int gi_trove_vectors(const gi_trove *t, int jind, int igamma, int *nsize, int *nrec,
                     const double **v);

/* The header line the reference prints before the transitions
 * (cuda_host.cu:1165) and the transition lines of rows [first, first+count)
 * of `lines`, in the format of cuda_host.cu:1444-1473, into buf (cap bytes,
 * NUL-terminated). *len receives the length written or, with GI_ECAPACITY,
 * the length needed. */
int gi_trove_header(const gi_trove *t, char *buf, long long cap, long long *len);
int gi_trove_format(const gi_trove *t, const gi_lines *lines, long long first, long long count,
                    char *buf, long long cap, long long *len);

#ifdef __cplusplus
}
#endif
#endif
