/*
 * gap_oracle.h -- CPU FP64 restatement of libgap's energy/force/stress path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle the HIP evaluator is
 * tested against: only the tests, the smoke check and the benchmark's CPU-baseline
 * leg may load it.  The product path never calls it.
 *
 * Layout conventions (C, row-major):
 *   lat[3*k + c]   component c of lattice vector k   (Fortran LAT(k+1,c+1))
 *   pos[3*i + c]   Cartesian position of atom i      (Fortran POS(i+1,c+1))
 *   force[3*i + c] force on atom i, eV/A
 *   stress6        GPa, order xx yy zz xy yz zx, VASP sign (libgap output)
 *   mm[k*D + d]    sparse point k, descriptor column d (raw, unstandardised)
 */
#ifndef GAP_ORACLE_H
#define GAP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nspecies;
    const int *z;          /* [nspecies] atomic numbers            */
    const double *w;       /* [nspecies] wACSF weights             */
    int nsf;               /* symmetry-function lines              */
    const int *type;       /* [nsf] 1..4                            */
    const double *alpha;   /* [nsf] eta (types 1,2,4) or Rs (type 3) */
    const double *cut;     /* [nsf] per-line cutoff                 */
    const double *zeta;    /* [nsf] angular exponent, NULL => 1     */
    int des_len;           /* D = 2*nsf                             */
    int nsparse;           /* M                                     */
    const double *theta;   /* [D]                                   */
    const double *mm;      /* [M*D]                                 */
    const double *coeff;   /* [M]                                   */
} orc_model;

/* Weight of each atom from the model header; returns the number of atoms
 * whose species is not listed (their weight is set to 0). */
int orc_species_weights(const orc_model *m, int na, const int *species, double *w);

/* Brute-force periodic neighbour list exactly as FGAP_CALC builds it.
 * nb[(i*max_nb + s)*6 + 0..5] = image xyz, distance, weight, atom index (0-based).
 * Returns 0, or -1 when an atom has more than max_nb neighbours. */
int orc_neighbors_ref(int na, const double *lat, const double *pos, const double *w,
                      double rcut, int max_nb, double *nb, int *cnt);

/* CAR2ACSF restated: features xx[i*D + d]; optional dense derivatives
 * dxdy[((i*na + m)*3 + c)*D + d] = d xx_i[d] / d R_{m,c} and
 * strs[((i*3 + a)*3 + c)*D + d]. */
void orc_car2acsf_dense(const orc_model *m, int na, int max_nb, const double *pos,
                        const double *nb, const int *cnt, double *xx, double *dxdy,
                        double *strs, int lgrad);

/* FGAP_CALC restated step by step (dense dxdy; O(N^2 D) memory).
 * Optional outputs (may be NULL): xx[na*D], dedg[na*D], eatom[na].
 * Returns 0, -1 (neighbour cap), -2 (unlisted species), -3 (bad model). */
int orc_gap_calc_faithful(const orc_model *m, int na, const int *species, const double *lat,
                          const double *pos, double rcut, int lgrad, double *ene,
                          double *force, double *stress6, double *variance, double *xx,
                          double *dedg, double *eatom);

/* Same quantities from a cell list and a per-atom fused pass whose backward
 * works on distance derivatives (no dense dxdy).  nthreads <= 0: OpenMP default. */
int orc_gap_calc_sparse(const orc_model *m, int na, const int *species, const double *lat,
                        const double *pos, double rcut, int lgrad, int nthreads, double *ene,
                        double *force, double *stress6, double *xx, double *dedg,
                        double *eatom);

/* Same as orc_gap_calc_sparse but only for centre atoms [i0, i1): the energy,
 * forces and virial are the contributions of those centres.  Used to time a
 * bounded sample of a large structure. */
int orc_gap_calc_sparse_range(const orc_model *m, int na, const int *species,
                              const double *lat, const double *pos, double rcut, int lgrad,
                              int nthreads, int i0, int i1, double *ene, double *force,
                              double *stress6);

/* Descriptor matrix only (sparse neighbour search), xx[i*D + d]. */
int orc_descriptors(const orc_model *m, int na, const int *species, const double *lat,
                    const double *pos, double rcut, int nthreads, double *xx);

/* Cell-list neighbour list in canonical form: for every atom i, the sorted
 * list of (j, n1, n2, n3) with |R_j + n.L - R_i| <= rcut, (j,n) != (i,0),
 * where R are the positions as given (no wrapping).  Returns total count;
 * writes at most cap entries of 4 ints plus the distance. */
long orc_neighbor_set(int na, const double *lat, const double *pos, double rcut, long cap,
                      int *out4, double *dist, int *count_per_atom);

#ifdef __cplusplus
}
#endif
#endif
