/*
 * FP64 CPU oracle for the line-intensity hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference
 * algorithm (ahmed-f-alrefaie/gpu-intensities) used as the checker for the
 * HIP path and as the CPU baseline. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it. The product library never links
 * or calls it.
 *
 * Every function cites the reference file:line it restates. Pinning of each
 * function (which independent check fixes it) is listed in oracle.c.
 */
#ifndef GI_ORACLE_H
#define GI_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* Basis for one J, the fields of TO_bset_contrT (trove_objects.h:49-72) that
 * the path reads, after correlate_index (trove_functions.cpp:680-858) and
 * compute_ijterms (trove_functions.cpp:927-966). */
typedef struct or_bset {
    int jval;
    int Maxsymcoeffs;          /* number of cases                          */
    int Maxcontracts;          /* dimen: number of primitive functions      */
    int mat_size;              /* third extent of repres                    */
    const int *icontr2icase;   /* [2*dimen]: icase at r, ilambda at r+dimen */
    const int *iroot_correlat_j0; /* [dimen]: J=0 (vibrational) index v(r) */
    const int *k;              /* [dimen]                                   */
    const int *ktau;           /* [dimen]: 2k+tau                           */
    const int *N;              /* [nrep*Maxsymcoeffs]: irr[g].N[c] at g*Maxsymcoeffs+c */
    const int *Ntotal;         /* [nrep]                                    */
    const int *ijterms;        /* [nrep*Maxsymcoeffs]: c + g*Maxsymcoeffs   */
    const double *const *repres; /* [nrep] -> Ntotal[g]*degen[g]*mat_size   */
    const int *nsize;          /* [nrep]: eigenvector length per symmetry   */
} or_bset;

/* One energy level, the fields of TO_PTeigen (trove_objects.h:74-88) the path reads. */
typedef struct or_level {
    double energy;
    int jind;          /* 0 or 1: index into jvals                        */
    int jval;
    int igamma;        /* 0-based symmetry                                */
    int ndeg;
    int ilevel;        /* 0-based level number within its (J, gamma) block */
    int irec0;         /* 0-based eigenvector record (column in the block) */
    const int *quanta; /* [nmodes+1]                                      */
} or_level;

/* Job: FintensityJob (fields.h:48-82) reduced to what the path reads. */
typedef struct or_job {
    int nrepres;
    const int *degen;
    int sym_type;             /* 1 C2v(M), 2 C3v(M), 3 D3h(M) (fields.h:9-11) */
    int nmodes;
    double zpe, temperature, q_stat;
    double freq_window[2], erange_lower[2], erange_upper[2];
    int jvals[2];
    const double *gns;
    const int *isym_pairs;
    const int *igamma_pair;
    int reduced;
    const int *quanta_lower;  /* [2*nmodes]: lo at i, hi at nmodes+i */
    const int *quanta_upper;
    double thresh_linestrength;  /* S >= this to be emitted (0: all)  */
    double thresh_intensity;     /* I >= this to be emitted (0: all)  */
    const or_bset *bset[2];      /* basis for jvals[0], jvals[1]      */
    int nvib;                    /* J=0 Maxcontracts                  */
    const double *dipole_flipped;/* [3*nvib*nvib]: mu_a(vF,vI) at vF + vI*nvib + a*nvib^2 */
    int jmax;
    const double *threej;        /* (jmax+1)^2*9, layout of precompute_threej */
    int nlevels;
    const or_level *levels;      /* energy sorted                     */
    const double *const *vectors;/* [jind*nrepres+g] -> nsize*nrec, record irec0 at irec0*nsize */
} or_job;

/* Output of or_run: one row per emitted transition, in the reference's
 * order (lower level by energy, then upper level by energy). */
typedef struct or_lines {
    long long n;
    int *ilevelI, *ilevelF;   /* indices into job->levels */
    double *nu, *S, *A_einst, *I;
    double *amp;              /* [n*4]: A[idegI + idegF*2] (cuda_host.cu:1467-1471 order) */
    double *amp_abs;          /* [n*4]: sum_r |c_F[r]*h[r]| per component (condition floor) */
} or_lines;

double or_fakt(double a);
double or_three_j(int j1, int j2, int j3, int k1, int k2, int k3);
void   or_precompute_threej(int jmax, double *threej);
void   or_correlate(const or_bset *b, const int *degen, int nrep, int ideg, int igamma,
                    const double *vec, double *out);
void   or_half_ls(const or_bset *bI, const or_bset *bF, const double *dipole_flipped, int nvib,
                  const double *threej, int jmax, const double *vec, double *half_ls, int noskip);
void   or_constants(double *out7);
int    or_energy_filter_lower(const or_job *job, int J, double energy, const int *quanta);
int    or_energy_filter_upper(const or_job *job, int J, double energy, const int *quanta);
int    or_intensity_filter(const or_job *job, int jI, int jF, double eI, double eF,
                           int gI, int gF, const int *qI, const int *qF);
int    or_degeneracy_filter(const or_job *job, int gI, int gF, int idegI, int idegF);
void   or_find_igamma_pair(int nrep, const int *isym_pairs, int *igamma_pair);
/* Runs the lower-level loop over the given lower levels (indices into
 * job->levels; NULL means all). nthreads <= 0 uses the OpenMP default.
 * Returns 0 on success. Free the result with or_free_lines. */
int    or_run(const or_job *job, int nlower, const int *lower_ids, int noskip, int nthreads,
              or_lines *out);
void   or_free_lines(or_lines *out);
int    or_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
