/*
 * gapx.h -- C ABI of the MI355X (gfx950) GAP energy/force/stress evaluator.
 *
 * Every entry point takes plain pointers and sizes.  Host arrays belong to the
 * caller; device memory belongs to the handle that allocated it.  A handle is
 * not thread-safe; separate handles are independent.  Every function returns
 * GAPX_OK (0) or a GAPX_ERR_* code; gapx_last_error() gives the message of
 * the calling thread's last failure.
 *
 * Units and conventions are libgap's (gappy/README.md:43-45,
 * gappy/libgap/gap_calc.f90:1-46): positions and cell in Angstrom, energy in
 * eV, forces in eV/Angstrom, stress in GPa in the order xx yy zz xy yz zx with
 * VASP's sign (positive = compressive load on the cell, the negative of ASE's
 * stress).  Lattice vectors are the rows of LAT (gap_calc.f90:98).
 *
 * Array layouts:
 *   GAPX_LAYOUT_C (default)  lat[3*k + c] = component c of vector k,
 *                            pos[3*i + c], force[3*i + c]
 *   GAPX_LAYOUT_F            Fortran LAT(3,3), POS(NA,3), FORCE(NA,3)
 *                            (column-major: pos[i + NA*c])
 */
#ifndef GAPX_H
#define GAPX_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define GAPX_OK 0
#define GAPX_ERR_ARG 1        /* invalid argument or size                      */
#define GAPX_ERR_FILE 2       /* model file missing or unreadable              */
#define GAPX_ERR_MODEL 3      /* inconsistent model (des_len != 2*nsf, ...)     */
#define GAPX_ERR_SPECIES 4    /* atomic number not listed in the model header   */
#define GAPX_ERR_NEIGHBORS 5  /* neighbour count above the configured cap       */
#define GAPX_ERR_HIP 6        /* HIP runtime failure                           */
#define GAPX_ERR_NOMEM 7      /* device or host allocation failed              */
#define GAPX_ERR_CAP 8        /* FGAP_READ fixed-size caps exceeded             */

/* ---- flags -------------------------------------------------------------- */
#define GAPX_GRAD 1           /* lgrad: forces and stress (else energy only)    */
#define GAPX_MIXED 2          /* FP32 descriptors + FP16 hi/lo MFMA contraction */
#define GAPX_LAYOUT_F 4       /* Fortran column-major lat/pos/force             */
#define GAPX_REF_NEIGHBOR_CAP 8 /* fail above 1000 neighbours like the reference */
#define GAPX_BF16 16          /* with GAPX_MIXED: BF16 hi/lo instead of FP16     */
#define GAPX_VARIANCE 64      /* gapx_calc: compute VARIANCE (else 0 as FGAP_CALC) */

typedef struct gapx_model gapx_model;
typedef struct gapx_ctx gapx_ctx;
typedef struct gapx_mlp gapx_mlp;

int gapx_version(void);
/* Measured vector-FMA throughput in TFLOP/s (no counterpart in the reference;
 * a measurement aid for the benchmark): tflops[0] FP32 FMA without packed
 * instructions (as the kernels are compiled), tflops[1] FP64 FMA.
 * Register-only microkernels on the current device, about 20 ms in all. */
int gapx_probe_fma_rates(double *tflops);
const char *gapx_last_error(void);
/* Number of visible GPUs (0 when no GPU or no driver). */
int gapx_device_count(int *n);

/* ---- model ----------------------------------------------------------------
 * gapx_model_create replaces the model half of FGAP_CALC's inputs
 * (gap_calc.f90:1-46: theta, MM, coeff) together with the header FGAP_CALC
 * and CAR2ACSF re-read from ./gap_parameters on every call
 * (gap_calc.f90:75-83 species weights, wacsf.f90:40-56 SF table).
 * mm_layout: 0 = row-major [nsparse][des_len]; 1 = Fortran MM(nsparse,des_len)
 * with leading dimension ld_mm (>= nsparse).  sf_zeta may be NULL (all 1).
 * The model is uploaded once and stays resident on the current device. */
int gapx_model_create(int nspecies, const int *z, const double *w, int nsf, const int *sf_type,
                      const double *sf_alpha, const double *sf_cut, const double *sf_zeta,
                      int des_len, int nsparse, const double *theta, const double *mm,
                      int mm_layout, int ld_mm, const double *coeff, gapx_model **out);

/* Parse a gap_parameters file (format of gap_calc.f90:303-364 and
 * wacsf.f90:40-56, read with Fortran list-directed rules) with no fixed caps;
 * replaces FGAP_READ (gap_calc.f90:303) plus the header reads. */
int gapx_model_read(const char *path, gapx_model **out);

int gapx_model_dims(const gapx_model *m, int *nspecies, int *nsf, int *des_len, int *nsparse);

/* Copy the model's GPR arrays back (FGAP_READ outputs).  mm_layout/ld_mm as
 * in gapx_model_create.  Any pointer may be NULL. */
int gapx_model_arrays(const gapx_model *m, double *theta, double *mm, int mm_layout, int ld_mm,
                      double *coeff);

/* Model header (species table and SF lines); arrays sized by gapx_model_dims. */
int gapx_model_header(const gapx_model *m, int *z, double *w, int *sf_type, double *sf_alpha,
                      double *sf_cut, double *sf_zeta);

void gapx_model_free(gapx_model *m);

/* How the mixed path treats this model (no reference counterpart; the
 * reference computes in FP64 only).  The energy E_i = sum_k W_ik is a sum of
 * terms W_ik = K_ik c_k that can cancel: its FP32 error grows with
 * A_i = sum_k |W_ik|.  cancellation = the largest A_j over up to 64 of the
 * model's own sparse points j (FP64, at load); mixed_contraction_fp64 = 1
 * when A * 2^-21 exceeds 1e-4 eV or a scaled sparse-point component exceeds
 * 2^14, and then GAPX_MIXED runs keep FP32 descriptors but contract in FP64.
 * The FP32 features: a relative error u of feature d moves E_j by about
 * u |x_jd / theta_d| |dE_j/dxt_d|; feature_sensitivity = the largest sum of
 * these over d at the same probed points (u = 1), and mixed_runs_fp64 = 1
 * when feature_sensitivity * 2^-24 exceeds 1e-4 eV: GAPX_MIXED runs of the
 * model are then FP64 runs (features, contraction and backward; results
 * equal the FP64 path's).  Any pointer may be NULL. */
int gapx_model_precision(const gapx_model *m, double *cancellation, int *mixed_contraction_fp64,
                         double *feature_sensitivity, int *mixed_runs_fp64);

/* Model I/O beyond FGAP_READ (gap_calc.f90:303-364 reads the text once per
 * call and keeps no cache).
 * gapx_model_write: the model as a gap_parameters text file in the layout
 * FGAP_READ and CAR2ACSF read (an SF line gets a 4th column zeta only when
 * some zeta differs from 1), 17 significant digits, so reading it back gives
 * the same doubles.
 * gapx_model_save / gapx_model_load: the same content as one binary file with
 * a checksum; no text parsing on load.
 * gapx_model_read_cached: gapx_model_read of a text file through a binary
 * copy (cache_path, or path + ".gapxbin" when NULL) that records the text
 * file's size and mtime; a missing, stale or damaged copy is rebuilt from the
 * text (failure to write it is not an error). */
int gapx_model_write(const gapx_model *m, const char *path);
int gapx_model_save(const gapx_model *m, const char *path);
int gapx_model_load(const char *path, gapx_model **out);
int gapx_model_read_cached(const char *path, const char *cache_path, gapx_model **out);

/* Size and modification time of a file, for callers that cache models keyed
 * on the file (the reference re-reads ./gap_parameters on every call). */
int gapx_file_stamp(const char *path, long long *size, long long *mtime_ns);

/* Q of the predictive variance (FGAP_CALC's QMM argument, gap_calc.f90:38,
 * which FGAP_READ leaves zero, gap_calc.f90:355-358): nsparse x nsparse,
 * layout 0 = row-major, 1 = Fortran QMM(nsparse,nsparse) with leading
 * dimension ld (>= nsparse).  NULL clears it (Q = 0).  Kept on the device. */
int gapx_model_set_qmm(gapx_model *m, const double *qmm, int layout, int ld);

/* ---- one-shot evaluation ---------------------------------------------------
 * FGAP_CALC (gap_calc.f90:1-301): energy, forces, stress and variance of one
 * periodic structure.  The variance is 0 as in the reference
 * (gap_calc.f90:206) unless flags has GAPX_VARIANCE: then it is the mean
 * over atoms of the predictive variance (see gapx_ctx_variance).  rcut is the
 * neighbour cutoff (per-line cutoffs above it are clipped by the neighbour
 * list, as in the reference).  Uses a workspace cached inside the model
 * handle. */
int gapx_calc(gapx_model *m, int na, const int *species, const double *lat, const double *pos,
              double rcut, int flags, double *ene, double *force, double *stress6,
              double *variance);

/* Batched FGAP_CALC over nstruct independent structures packed back to back
 * (natoms[s] atoms each; species/pos concatenated; lat is nstruct*9).
 * Always GAPX_LAYOUT_C.  ene[nstruct], force[3*sum(natoms)], stress6[6*nstruct]. */
int gapx_calc_batch(gapx_model *m, int nstruct, const int *natoms, const int *species,
                    const double *lat, const double *pos, double rcut, int flags, double *ene,
                    double *force, double *stress6);

/* Descriptors only (CAR2ACSF's xx, wacsf.f90:2-12; FP64): x[na*des_len],
 * row = atom, column order = the SF table's plain features then weighted. */
int gapx_descriptors(gapx_model *m, int na, const int *species, const double *lat,
                     const double *pos, double rcut, int flags, double *x);

/* CAR2ACSF (gappy/libgap/wacsf.f90:2-12): wACSF descriptors and their dense
 * derivatives from a neighbour table the caller built, in the reference's
 * Fortran layouts.  pos(NA, 3); neighbor(NA, max_neighbor, 6) with, per slot,
 * the image position x y z, the distance, the weight and the 1-based index of
 * the neighbour atom as a real (the table FGAP_CALC builds,
 * gap_calc.f90:89-120); neighbor_count(NA).  Outputs (FP64): xx(nf, NA);
 * dxdy(nf, NA, NA, 3) = d xx(f, i) / d pos(n, c); strs(3, 3, nf, NA) =
 * sum over the neighbour terms of delta_a d xx(f, i) / d pos(j, c).  The SF
 * table is the model's (only its header is used); nf must be its des_len.
 * With lgrad = 0, or for a NULL pointer, dxdy and strs are not computed (a
 * non-NULL one is zeroed, as the reference zeroes both).  A distance of 0 or
 * an index outside [1, NA] is GAPX_ERR_ARG. */
int gapx_car2acsf(gapx_model *m, int na, int max_neighbor, int nf, const double *pos, const double *neighbor,
                  const int *neighbor_count, double *xx, double *dxdy, double *strs, int lgrad);

/* FGET_BOND (gappy/libgap/get_bond.f90:4-117): min(10, shortest distance
 * within rcut between any atom and any periodic image of any atom, the
 * atom's own zero image excluded), over the reference's image range.
 * lat/pos in the layout given by flags (GAPX_LAYOUT_F or C). */
int gapx_min_distance(int na, const double *lat, const double *pos, double rcut, int flags, double *dmin);

/* gapx_min_distance without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight screening loops, harnesses that forbid hipMalloc).  Same
 * arguments and result as gapx_min_distance (bit for bit); the workspace is
 * the library's fixed 16 MiB device array, one per GPU, kept until the
 * process ends.  Inputs whose workspace (staged positions, bins, binned
 * copies: 148 bytes per atom plus 8 to 12 per bin) exceeds it are refused
 * with GAPX_ERR_ARG before any GPU work, roughly above 110 000 atoms.  Every call
 * on one GPU works in that GPU's single array: calls that overlap on one GPU
 * (from several host threads) race; serialise them. */
int gapx_min_distance_noalloc(int na, const double *lat, const double *pos, double rcut, int flags,
                              double *dmin);

/* ---- device-resident contexts ----------------------------------------------
 * A context owns two HIP streams (its work stream, and a second one joined
 * inside each run) and device buffers for a batch of structures.
 * Load once, then run repeatedly on the inputs already in HBM; results stay
 * on the device until fetched. */
int gapx_ctx_create(gapx_model *m, gapx_ctx **out);
int gapx_ctx_load(gapx_ctx *c, int nstruct, const int *natoms, const int *species,
                  const double *lat, const double *pos, double rcut);
/* New positions (and cells, may be NULL) for the loaded batch, same sizes. */
int gapx_ctx_update(gapx_ctx *c, const double *lat, const double *pos);
/* Enqueue one full evaluation on the context's stream (asynchronous). */
int gapx_ctx_run(gapx_ctx *c, int flags);
int gapx_ctx_sync(gapx_ctx *c);
/* Fetch results of the last run (synchronises).  Any pointer may be NULL.
 * eatom gives the per-atom (local) energies. */
int gapx_ctx_results(gapx_ctx *c, double *ene, double *force, double *stress6, double *eatom);
/* Per-atom descriptors (raw, FP64 or FP32 run) and dE/dx of the last run. */
int gapx_ctx_features(gapx_ctx *c, double *x, double *dedx);
/* Predictive variance of the last run's structures, the formula FGAP_CALC
 * keeps commented out (gap_calc.f90:205-210): per atom
 *   sigma2_i = 1 - k_i Q k_i^T,  k_ik = exp(-sum_d ((x_id - MM_kd)/theta_d)^2 / 2)
 * (the covariance of gap_calc.f90:270-287, delta = 1), Q from
 * gapx_model_set_qmm (sigma2_i = 1 while none is set), and var[s] the mean
 * over structure s's atoms.  FP64 (a mixed run's FP32 features are widened).
 * var[nstruct], sigma2[sum natoms]; either may be NULL.  Synchronises. */
int gapx_ctx_variance(gapx_ctx *c, double *var, double *sigma2);
/* Stage timing with HIP events on the context's stream: enable, then read
 * the summed milliseconds and launch counts per stage since the last reset. */
int gapx_ctx_timing(gapx_ctx *c, int enable);
int gapx_ctx_stage_count(gapx_ctx *c, int *n);
int gapx_ctx_stage_info(gapx_ctx *c, int k, const char **name, double *ms, int *launches);
/* Capture gapx_ctx_run(flags) into a HIP graph and replay it on later runs
 * with the same flags, atom count and list capacities (on = 1), or go back to
 * eager launches (on = 0).  The first run after a change is captured after
 * one eager run; runs with timing enabled are always eager.  No reference
 * counterpart (the reference is a CPU library). */
int gapx_ctx_use_graph(gapx_ctx *c, int on);
/* Relax every loaded structure at fixed cell with FIRE (ASE's FIRE defaults:
 * dt 0.1, dtmax 1.0, Nmin 5, finc 1.1, fdec 0.5, astart 0.1, fa 0.99) until
 * max |F_i| <= fmax (eV/A) for every structure or max_steps steps; a step
 * moves each atom by at most maxstep (A, the norm over the structure).
 * Positions are updated on the device; gapx_ctx_results then gives the
 * energies and forces at the final positions, gapx_ctx_positions the
 * positions (C layout).  Replaces the reference's per-structure relaxation
 * processes (gappy/example/ASE-GAPPY/run_opt.py:108-109 with LBFGS, driven
 * by tools/cgg2.py:100-137); the optimiser differs (FIRE, not LBFGS). */
int gapx_ctx_relax(gapx_ctx *c, int flags, int max_steps, double fmax, double maxstep, int *steps_done,
                   int *n_converged);
int gapx_ctx_positions(gapx_ctx *c, double *pos);
/* Relax every loaded structure at fixed cell with LBFGS as ASE 3.18.1 runs
 * it, the reference's optimiser (gappy/example/ASE-GAPPY/run_opt.py:108-109:
 * LBFGS(atoms) at its defaults, run(fmax = -EDIFFG, steps = NSW)): history
 * of `memory` steps (ASE: 100), initial inverse Hessian 1 / alpha (ASE: 70),
 * no line search, each step scaled so that no atom moves more than maxstep
 * (ASE: 0.2 A); a structure stops when max_i |F_i|^2 < fmax^2.  At most
 * max_steps steps per structure; *steps_done = steps of the slowest
 * structure, *n_converged = structures converged at their final positions.
 * gapx_ctx_results / gapx_ctx_positions then give the final state. */
int gapx_ctx_relax_lbfgs(gapx_ctx *c, int flags, int max_steps, double fmax, double maxstep, int memory,
                         double alpha, int *steps_done, int *n_converged);

/* Batched NVT molecular dynamics on the device: nsteps of ASE's NVTBerendsen
 * step (ASE 3.18.1), which the reference's MD driver runs for one structure
 * per process (gappy/example/ASE-GAPPY/run_md.py:10-21, timestep POTIM fs,
 * temperature TEBEG K, taut 500 fs), for every loaded structure at once:
 * Berendsen scale s = sqrt(1 + (T0/T - 1) dt/taut) clipped to [0.9, 1.1]
 * (taut <= 0: no scaling), p <- s p + dt/2 F, total momentum removed when
 * fixcm, x <- x + dt p/m, forces at x, p <- p + dt/2 F.  ASE units: dt in
 * ASE time units (1 fs = 0.09822694788464063), masses[N] in amu, momenta[3N]
 * in amu A per time unit (NULL: start at rest), temperature in K with
 * kB = 8.6173303e-5 eV/K.  log (may be NULL): [nsteps+1][nstruct][3] rows of
 * (E_pot eV, E_kin eV, T K), row 0 at the start.  Positions and momenta stay
 * on the device (gapx_ctx_positions, gapx_ctx_momenta); gapx_ctx_results
 * gives energies and forces at the final positions. */
int gapx_ctx_md(gapx_ctx *c, int flags, int nsteps, double dt, double temperature, double taut, int fixcm,
                const double *masses, const double *momenta, double *log);
int gapx_ctx_momenta(gapx_ctx *c, double *p);
/* Pick launch shapes for runs with these flags (wavefronts per atom in the
 * descriptor kernels, contraction row tile and workgroup count, and whether a
 * large batch runs as two overlapped halves) by timing the loaded batch;
 * results are cached per process for the shape class unless force != 0.
 * No reference counterpart. */
int gapx_ctx_autotune(gapx_ctx *c, int flags, int force);
int gapx_ctx_get_tuning(gapx_ctx *c, int flags, int *waves_fwd, int *waves_bwd, int *gpr_rows, int *gpr_blocks,
                        int *halves);
int gapx_ctx_set_tuning(gapx_ctx *c, int flags, int waves_fwd, int waves_bwd, int gpr_rows, int gpr_blocks,
                        int halves);
/* Work counts of the last run: (centre, neighbour) pairs within rcut and
 * neighbour pairs j < k of a centre within the largest angular cutoff. */
int gapx_ctx_counts(gapx_ctx *c, long long *pairs, long long *triplets);
/* Bytes of device memory held by the context. */
int gapx_ctx_device_bytes(gapx_ctx *c, long long *bytes);
void gapx_ctx_free(gapx_ctx *c);

/* ---------------------------------------------------------------------------
 * Per-atom MLP surrogate (no reference counterpart): 128 -> 128 -> 128 -> 1,
 * SiLU, on the standardised descriptors xt = x theta_inv - mu of a model with
 * 128 descriptor columns.  FP32 weights, BF16 MFMA products (FP32
 * accumulation); GAPX_MLP_SPLIT uses hi + lo BF16 operands (FP32-level
 * products).  Trained on GAP per-atom energies e_i and per-atom gradients
 * G_i = dE_i/dxt_i, which fix the forces through the descriptor Jacobian:
 *   L = cE/(2n) sum_i (e_i - et_i)^2 + cG/(2n) sum_i |G_i - Gt_i|^2.
 * A handle is not thread-safe. */
#define GAPX_MLP_SPLIT 1
/* run flag of gapx_ctx_run_surrogate: hi + lo products in the MLP */
#define GAPX_MLP_EXACT 32
int gapx_mlp_param_count(int *n);
int gapx_mlp_create(unsigned long long seed, gapx_mlp **out);
void gapx_mlp_free(gapx_mlp *m);
/* Flat FP32 parameters: W1[128][128] (row = input), b1[128], W2[128][128],
 * b2[128], w3[128], b3.  Setting them resets the optimiser state. */
int gapx_mlp_get_params(gapx_mlp *m, float *params);
int gapx_mlp_set_params(gapx_mlp *m, const float *params);
/* e[n] and, if G is not NULL, G[n][128] = de/dxt for xt[n][128] (host). */
int gapx_mlp_forward(gapx_mlp *m, int n, const float *xt, int flags, float *e, float *G);
/* Loss sums (sum (e - et)^2, sum |G - Gt|^2) and dL/dparams for one batch. */
int gapx_mlp_loss_grad(gapx_mlp *m, int n, const float *xt, const float *e_t, const float *G_t, float cE, float cG,
                       int flags, double *sums, float *grad);
/* Device-resident training set: append rows (host arrays), clear, count. */
int gapx_mlp_data_add(gapx_mlp *m, int n, const float *xt, const float *e_t, const float *G_t);
int gapx_mlp_data_clear(gapx_mlp *m);
int gapx_mlp_data_count(gapx_mlp *m, long long *n);
/* `steps` Adam steps on uniformly sampled batches of `batch` rows; sums gets
 * the last batch's two sums of squares. */
int gapx_mlp_train(gapx_mlp *m, int steps, int batch, float lr, float cE, float cG, unsigned long long seed,
                   int flags, double *sums);
/* Append the context's last GAP run (GAPX_GRAD) to the training set: xt, the
 * per-atom energies and dE_i/dxt_i, device to device. */
int gapx_ctx_export_labels(gapx_ctx *c, gapx_mlp *m);
/* Evaluate the loaded batch with the surrogate in place of the GAP
 * contraction: neighbours, descriptors (FP64, or FP32 with GAPX_MIXED), MLP
 * energies and, with GAPX_GRAD, forces and stress through the descriptor
 * backward pass.  Results through gapx_ctx_results. */
int gapx_ctx_run_surrogate(gapx_ctx *c, gapx_mlp *m, int flags);

#ifdef __cplusplus
}
#endif
#endif
