/*
 * gap_oracle.c -- CPU FP64 restatement of libgap's energy/force/stress path.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the oracle the HIP evaluator is tested against.
 * Only the tests, the smoke check and the benchmark's CPU-baseline leg may
 * load it; the product path never does.
 *
 * Two modes:
 *   faithful -- follows gappy/libgap/gap_calc.f90 (FGAP_CALC, lines 1-301) and
 *               gappy/libgap/wacsf.f90 (CAR2ACSF, lines 2-796) step by step:
 *               brute-force image search, one sweep per symmetry-function
 *               line, dense d(feature)/dR, GET_COV, dE/dG, dense force and
 *               stress contractions, output reorder.
 *   sparse   -- the same quantities from a cell list and a fused per-atom
 *               pass; the backward pass contracts dE/dG with derivatives of
 *               each term with respect to the pair distances (r_ij, r_ik,
 *               r_jk), which is the formulation the GPU kernels use.
 *
 * Extension (no reference counterpart): an optional angular exponent zeta per
 * line.  The angular term is 2^(1-zeta) (1 + lambda cos)^zeta; at zeta = 1 it
 * is exactly the reference's (1 + lambda cos) (wacsf.f90:276, 623).
 */
#include "gap_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.141592654     /* truncated pi, wacsf.f90:31 and gap_calc.f90:59 */
#define ORC_GPA 6.24219e-3     /* eV/A^3 per GPa, gap_calc.f90:9 */
#define ORC_DELTA 1.0          /* GPR signal variance, gap_calc.f90:8 */

/* ------------------------------------------------------------------------ */
/* small helpers                                                            */
/* ------------------------------------------------------------------------ */

static double zeta_of(const orc_model *m, int s) { return m->zeta ? m->zeta[s] : 1.0; }

/* (1 + lambda*c)^zeta * 2^(1-zeta) and its derivative with respect to c. */
static void ang_power(double base, double lam, double zeta, double *p, double *dp) {
    if (zeta == 1.0) {
        *p = base;
        *dp = lam;
        return;
    }
    double pref = pow(2.0, 1.0 - zeta);
    double bz1;
    if (zeta == floor(zeta) && zeta > 1.0 && zeta < 64.0) {
        int n = (int)zeta - 1;
        bz1 = 1.0;
        for (int t = 0; t < n; ++t) bz1 *= base;
    } else {
        bz1 = pow(base, zeta - 1.0);
    }
    *p = pref * bz1 * base;
    *dp = pref * zeta * lam * bz1;
}

static double fcut(double r, double rc) { return 0.5 * (cos(ORC_PI * r / rc) + 1.0); }
static double dfcut(double r, double rc) { return 0.5 * (-sin(ORC_PI * r / rc)) * (ORC_PI / rc); }

static void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* Det() of gap_calc.f90:290-299 on Fortran LAT(3,3); lat is row-major so
 * LAT(r,c) = lat[3*(r-1) + (c-1)]. */
static double det_lat(const double *L) {
    return L[0] * (L[4] * L[8] - L[5] * L[7]) - L[1] * (L[3] * L[8] - L[5] * L[6]) +
           L[2] * (L[3] * L[7] - L[6] * L[4]);
}

int orc_species_weights(const orc_model *m, int na, const int *species, double *w) {
    /* gap_calc.f90:75-83: every header line assigns its weight to the atoms of
     * that Z; later lines win.  An unlisted Z is left undefined there; here it
     * is reported and set to 0. */
    int missing = 0;
    for (int j = 0; j < na; ++j) {
        int found = 0;
        w[j] = 0.0;
        for (int t = 0; t < m->nspecies; ++t)
            if (species[j] == m->z[t]) {
                w[j] = m->w[t];
                found = 1;
            }
        if (!found) ++missing;
    }
    return missing;
}

static int model_ok(const orc_model *m) {
    if (m->nsf <= 0 || m->des_len != 2 * m->nsf || m->nsparse <= 0) return 0;
    for (int s = 0; s < m->nsf; ++s)
        if (m->cut[s] <= 0.0) return 0;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* faithful mode                                                            */
/* ------------------------------------------------------------------------ */

int orc_neighbors_ref(int na, const double *lat, const double *pos, const double *w,
                      double rcut, int max_nb, double *nb, int *cnt) {
    /* gap_calc.f90:85-88 -- reciprocal vectors from the columns of LAT
     * (recipvector, :240-249), image counts n_k = ceil(rcut*|b_k|/pi/2). */
    double col[3][3], rc[3][3];
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 3; ++k) col[c][k] = lat[3 * k + c];
    cross3(col[1], col[2], rc[0]);
    cross3(col[2], col[0], rc[1]);
    cross3(col[0], col[1], rc[2]);
    double vol = fabs(col[0][0] * rc[0][0] + col[0][1] * rc[0][1] + col[0][2] * rc[0][2]);
    int nabc[3];
    for (int r = 0; r < 3; ++r) {
        /* row r of the reciprocal matrix whose columns are rc[0..2] */
        double v0 = rc[0][r] / vol * ORC_PI * 2.0;
        double v1 = rc[1][r] / vol * ORC_PI * 2.0;
        double v2 = rc[2][r] / vol * ORC_PI * 2.0;
        double len = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
        nabc[r] = (int)ceil(rcut * len / ORC_PI / 2.0);
    }
    /* gap_calc.f90:92-120 -- loop i, j, n1, n2, n3 in that order. */
    for (int i = 0; i < na; ++i) {
        cnt[i] = 0;
        for (int j = 0; j < na; ++j)
            for (int n1 = -nabc[0]; n1 <= nabc[0]; ++n1)
                for (int n2 = -nabc[1]; n2 <= nabc[1]; ++n2)
                    for (int n3 = -nabc[2]; n3 <= nabc[2]; ++n3) {
                        if (n1 == 0 && n2 == 0 && n3 == 0 && i == j) continue;
                        double xyz[3], dr[3];
                        for (int c = 0; c < 3; ++c) {
                            xyz[c] = pos[3 * j + c] + (double)n1 * lat[c] + (double)n2 * lat[3 + c] +
                                     (double)n3 * lat[6 + c];
                            dr[c] = pos[3 * i + c] - xyz[c];
                        }
                        double dis = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
                        if (dis > rcut) continue;
                        if (cnt[i] >= max_nb) return -1;
                        double *e = nb + ((size_t)i * max_nb + cnt[i]) * 6;
                        e[0] = xyz[0];
                        e[1] = xyz[1];
                        e[2] = xyz[2];
                        e[3] = dis;
                        e[4] = w[j];
                        e[5] = (double)j;
                        cnt[i]++;
                    }
    }
    return 0;
}

/* Accumulate one term's derivative into the dense arrays (wacsf.f90 pattern:
 * centre atom i, neighbour atoms n and m, stress from the neighbour parts). */
static inline void acc_dense(double *dxdy, int na, int D, int i, int d, int atom,
                             const double *grad, double scale) {
    for (int c = 0; c < 3; ++c) dxdy[(((size_t)i * na + atom) * 3 + c) * D + d] += grad[c] * scale;
}

void orc_car2acsf_dense(const orc_model *m, int na, int max_nb, const double *pos,
                        const double *nb, const int *cnt, double *xx, double *dxdy,
                        double *strs, int lgrad) {
    const int nsf = m->nsf, D = m->des_len;
    memset(xx, 0, sizeof(double) * (size_t)na * D);
    if (lgrad) {
        if (dxdy) memset(dxdy, 0, sizeof(double) * (size_t)na * na * 3 * D);
        if (strs) memset(strs, 0, sizeof(double) * (size_t)na * 9 * D);
    }
    for (int s = 0; s < nsf; ++s) {
        const int type = m->type[s];
        const double rc = m->cut[s];
        if (type == 1 || type == 3) {
            /* wacsf.f90:69-162 (type 1) and 436-527 (type 3) */
            const double alpha = (type == 1) ? m->alpha[s] : 4.0;
            const double rs = (type == 1) ? 0.0 : m->alpha[s];
            for (int i = 0; i < na; ++i)
                for (int a = 0; a < cnt[i]; ++a) {
                    const double *e = nb + ((size_t)i * max_nb + a) * 6;
                    double rij = e[3];
                    if (rij > rc) continue;
                    double wj = e[4];
                    int n = (int)e[5];
                    double fc = fcut(rij, rc);
                    double ex = (type == 1) ? exp(-alpha * rij * rij) : exp(-alpha * (rij - rs) * (rij - rs));
                    xx[(size_t)i * D + s] += ex * fc;
                    xx[(size_t)i * D + s + nsf] += ex * fc * wj;
                    if (!lgrad) continue;
                    double del[3], drj[3];
                    for (int c = 0; c < 3; ++c) {
                        del[c] = -(pos[3 * i + c] - e[c]);
                        drj[c] = del[c] / rij; /* d rij / d R_j */
                    }
                    double dfc = dfcut(rij, rc);
                    double t1 = (type == 1) ? -2.0 * alpha * rij * ex * fc : -2.0 * alpha * (rij - rs) * ex * fc;
                    double gi[3], gj[3];
                    for (int c = 0; c < 3; ++c) {
                        gj[c] = drj[c] * t1 + ex * dfc * drj[c];
                        gi[c] = -drj[c] * t1 + ex * dfc * (-drj[c]);
                    }
                    if (dxdy) {
                        acc_dense(dxdy, na, D, i, s, i, gi, 1.0);
                        acc_dense(dxdy, na, D, i, s + nsf, i, gi, wj);
                        acc_dense(dxdy, na, D, i, s, n, gj, 1.0);
                        acc_dense(dxdy, na, D, i, s + nsf, n, gj, wj);
                    }
                    if (strs)
                        for (int aa = 0; aa < 3; ++aa)
                            for (int c = 0; c < 3; ++c) {
                                strs[(((size_t)i * 3 + aa) * 3 + c) * D + s] += del[aa] * gj[c];
                                strs[(((size_t)i * 3 + aa) * 3 + c) * D + s + nsf] += del[aa] * gj[c] * wj;
                            }
                }
        } else if (type == 2 || type == 4) {
            /* wacsf.f90:169-432 (lambda = +1) and 534-791 (lambda = -1) */
            const double alpha = m->alpha[s];
            const double lam = (type == 2) ? 1.0 : -1.0;
            const double zeta = zeta_of(m, s);
            for (int i = 0; i < na; ++i)
                for (int ja = 0; ja < cnt[i]; ++ja) {
                    const double *ej = nb + ((size_t)i * max_nb + ja) * 6;
                    double rij = ej[3];
                    if (rij > rc) continue;
                    double fij = fcut(rij, rc), wj = ej[4];
                    int n = (int)ej[5];
                    for (int ka = ja + 1; ka < cnt[i]; ++ka) {
                        const double *ek = nb + ((size_t)i * max_nb + ka) * 6;
                        double rik = ek[3];
                        if (rik > rc) continue;
                        double fik = fcut(rik, rc), wk = ek[4];
                        int mm_ = (int)ek[5];
                        double djk[3];
                        for (int c = 0; c < 3; ++c) djk[c] = ej[c] - ek[c];
                        double rjk = sqrt(djk[0] * djk[0] + djk[1] * djk[1] + djk[2] * djk[2]);
                        if (rjk > rc) continue;
                        double fjk = fcut(rjk, rc);
                        double f = rjk * rjk - rij * rij - rik * rik;
                        double g = -2.0 * rij * rik;
                        double cth = f / g;
                        double P, dP;
                        ang_power(1.0 + lam * cth, lam, zeta, &P, &dP);
                        double ex = exp(-alpha * (rij * rij + rik * rik + rjk * rjk));
                        double F3 = fij * fik * fjk;
                        double term = P * ex * F3;
                        xx[(size_t)i * D + s] += term;
                        xx[(size_t)i * D + s + nsf] += term * wj * wk;
                        if (!lgrad) continue;
                        /* derivatives with respect to R_i, R_j, R_k (wacsf.f90:182-325) */
                        double dj[3], dk[3];
                        for (int c = 0; c < 3; ++c) {
                            dj[c] = -(pos[3 * i + c] - ej[c]);
                            dk[c] = -(pos[3 * i + c] - ek[c]);
                        }
                        double dfij = dfcut(rij, rc), dfik = dfcut(rik, rc), dfjk = dfcut(rjk, rc);
                        double grad[3][3]; /* [i,j,k][c] */
                        for (int c = 0; c < 3; ++c) {
                            double drij[3] = {-dj[c] / rij, dj[c] / rij, 0.0};
                            double drik[3] = {-dk[c] / rik, 0.0, dk[c] / rik};
                            double drjk[3] = {0.0, djk[c] / rjk, -djk[c] / rjk};
                            for (int X = 0; X < 3; ++X) {
                                double df = 2.0 * rjk * drjk[X] - 2.0 * rij * drij[X] - 2.0 * rik * drik[X];
                                double dg = -2.0 * (drij[X] * rik + rij * drik[X]);
                                double dcos = (df * g - f * dg) / (g * g);
                                double dex = -alpha * 2.0 * ex * (rij * drij[X] + rik * drik[X] + rjk * drjk[X]);
                                double dF = dfij * drij[X] * fik * fjk + fij * dfik * drik[X] * fjk +
                                            fij * fik * dfjk * drjk[X];
                                grad[X][c] = dP * dcos * ex * F3 + P * dex * F3 + P * ex * dF;
                            }
                        }
                        double wjk = wj * wk;
                        if (dxdy) {
                            acc_dense(dxdy, na, D, i, s, i, grad[0], 1.0);
                            acc_dense(dxdy, na, D, i, s, n, grad[1], 1.0);
                            acc_dense(dxdy, na, D, i, s, mm_, grad[2], 1.0);
                            acc_dense(dxdy, na, D, i, s + nsf, i, grad[0], wjk);
                            acc_dense(dxdy, na, D, i, s + nsf, n, grad[1], wjk);
                            acc_dense(dxdy, na, D, i, s + nsf, mm_, grad[2], wjk);
                        }
                        if (strs)
                            for (int aa = 0; aa < 3; ++aa)
                                for (int c = 0; c < 3; ++c) {
                                    double v = dj[aa] * grad[1][c] + dk[aa] * grad[2][c];
                                    strs[(((size_t)i * 3 + aa) * 3 + c) * D + s] += v;
                                    strs[(((size_t)i * 3 + aa) * 3 + c) * D + s + nsf] += v * wjk;
                                }
                    }
                }
        }
        /* any other type: feature stays 0 (wacsf.f90:792-794) */
    }
}

int orc_gap_calc_faithful(const orc_model *m, int na, const int *species, const double *lat,
                          const double *pos, double rcut, int lgrad, double *ene,
                          double *force, double *stress6, double *variance, double *xx_out,
                          double *dedg_out, double *eatom_out) {
    if (!model_ok(m)) return -3;
    const int D = m->des_len, M = m->nsparse;
    const int max_nb = 1000; /* gap_calc.f90:68 */
    double *w = (double *)malloc(sizeof(double) * na);
    if (orc_species_weights(m, na, species, w)) {
        free(w);
        return -2;
    }
    double *nb = (double *)calloc((size_t)na * max_nb * 6, sizeof(double));
    int *cnt = (int *)calloc(na, sizeof(int));
    if (orc_neighbors_ref(na, lat, pos, w, rcut, max_nb, nb, cnt)) {
        free(w);
        free(nb);
        free(cnt);
        return -1;
    }
    double *xx = (double *)malloc(sizeof(double) * (size_t)na * D);
    double *dxdy = lgrad ? (double *)malloc(sizeof(double) * (size_t)na * na * 3 * D) : NULL;
    double *strs = lgrad ? (double *)malloc(sizeof(double) * (size_t)na * 9 * D) : NULL;
    orc_car2acsf_dense(m, na, max_nb, pos, nb, cnt, xx, dxdy, strs, lgrad);
    free(nb);
    free(cnt);
    free(w);

    /* GET_COV (gap_calc.f90:268-288) and energies (:152-154) */
    double *ckm = (double *)malloc(sizeof(double) * (size_t)na * M);
    for (int i = 0; i < na; ++i)
        for (int k = 0; k < M; ++k) {
            double t = 0.0;
            for (int d = 0; d < D; ++d) {
                double u = (xx[(size_t)i * D + d] - m->mm[(size_t)k * D + d]) / m->theta[d];
                t += u * u;
            }
            ckm[(size_t)i * M + k] = ORC_DELTA * exp(-0.5 * t);
        }
    double etot = 0.0;
    for (int i = 0; i < na; ++i) {
        double e = 0.0;
        for (int k = 0; k < M; ++k) e += ckm[(size_t)i * M + k] * m->coeff[k];
        if (eatom_out) eatom_out[i] = e;
        etot += e;
    }
    /* dE/dG (gap_calc.f90:157-166) */
    double *dedg = (double *)calloc((size_t)na * D, sizeof(double));
    for (int i = 0; i < na; ++i)
        for (int d = 0; d < D; ++d) {
            double acc = 0.0;
            double th2 = m->theta[d] * m->theta[d];
            for (int k = 0; k < M; ++k)
                acc -= (xx[(size_t)i * D + d] - m->mm[(size_t)k * D + d]) / th2 * ckm[(size_t)i * M + k] *
                       m->coeff[k];
            dedg[(size_t)i * D + d] = acc;
        }
    free(ckm);
    /* forces (gap_calc.f90:172-185) and stress (:188-203) */
    double sc[6] = {0, 0, 0, 0, 0, 0};
    for (int a = 0; a < na * 3; ++a) force[a] = 0.0;
    if (lgrad) {
        for (int mi = 0; mi < na; ++mi)
            for (int n = 0; n < na; ++n)
                for (int c = 0; c < 3; ++c) {
                    const double *dd = dxdy + (((size_t)n * na + mi) * 3 + c) * D;
                    double acc = 0.0;
                    for (int d = 0; d < D; ++d) acc -= dedg[(size_t)n * D + d] * dd[d];
                    force[3 * mi + c] += acc;
                }
        int k1 = 0;
        for (int a = 0; a < 3; ++a)
            for (int c = a; c < 3; ++c) {
                double acc = 0.0;
                for (int n = 0; n < na; ++n)
                    for (int d = 0; d < D; ++d)
                        acc -= dedg[(size_t)n * D + d] * strs[(((size_t)n * 3 + a) * 3 + c) * D + d];
                sc[k1++] = acc;
            }
    }
    double vol = fabs(det_lat(lat));
    for (int t = 0; t < 6; ++t) sc[t] = sc[t] * (1.0 / ORC_GPA) / vol;
    /* output order (gap_calc.f90:221-226): 11 22 33 12 23 13 */
    stress6[0] = sc[0];
    stress6[1] = sc[3];
    stress6[2] = sc[5];
    stress6[3] = sc[1];
    stress6[4] = sc[4];
    stress6[5] = sc[2];
    *ene = etot;
    if (variance) *variance = 0.0;
    if (xx_out) memcpy(xx_out, xx, sizeof(double) * (size_t)na * D);
    if (dedg_out) memcpy(dedg_out, dedg, sizeof(double) * (size_t)na * D);
    free(xx);
    free(dxdy);
    free(strs);
    free(dedg);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* cell list (sparse mode)                                                  */
/* ------------------------------------------------------------------------ */

typedef struct {
    int na;
    double L[9], inv[9];        /* rows = lattice vectors; inv = L^-1 */
    double *wpos;               /* wrapped positions [na*3] */
    int *shift;                 /* floor(frac) per atom [na*3] */
    int nb[3], S[3];            /* bins per axis, search half-width */
    int *bin_start, *bin_atoms; /* CSR by bin */
} cell_list;

typedef struct {
    int j;
    int n[3];                   /* image of j relative to the given positions */
    double d[3];                /* R_j + n.L - R_i */
    double r, w;
} nbr;

static void inv3(const double *A, double *B) {
    double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                 A[2] * (A[3] * A[7] - A[4] * A[6]);
    B[0] = (A[4] * A[8] - A[5] * A[7]) / det;
    B[1] = (A[2] * A[7] - A[1] * A[8]) / det;
    B[2] = (A[1] * A[5] - A[2] * A[4]) / det;
    B[3] = (A[5] * A[6] - A[3] * A[8]) / det;
    B[4] = (A[0] * A[8] - A[2] * A[6]) / det;
    B[5] = (A[2] * A[3] - A[0] * A[5]) / det;
    B[6] = (A[3] * A[7] - A[4] * A[6]) / det;
    B[7] = (A[1] * A[6] - A[0] * A[7]) / det;
    B[8] = (A[0] * A[4] - A[1] * A[3]) / det;
}

static void cell_build(cell_list *cl, int na, const double *lat, const double *pos, double rcut) {
    cl->na = na;
    memcpy(cl->L, lat, sizeof(double) * 9);
    inv3(lat, cl->inv);
    cl->wpos = (double *)malloc(sizeof(double) * 3 * (na > 0 ? na : 1));
    cl->shift = (int *)malloc(sizeof(int) * 3 * (na > 0 ? na : 1));
    /* interplanar spacing d_k = 1/|column k of L^-1| */
    for (int k = 0; k < 3; ++k) {
        double g = sqrt(cl->inv[k] * cl->inv[k] + cl->inv[3 + k] * cl->inv[3 + k] +
                        cl->inv[6 + k] * cl->inv[6 + k]);
        double dk = 1.0 / g;
        int nbk = (int)floor(dk / rcut);
        if (nbk < 1) nbk = 1;
        if (nbk > 64) nbk = 64;
        cl->nb[k] = nbk;
        cl->S[k] = (int)ceil(rcut * nbk / dk);
    }
    int nbins = cl->nb[0] * cl->nb[1] * cl->nb[2];
    int *bin = (int *)malloc(sizeof(int) * (na > 0 ? na : 1));
    cl->bin_start = (int *)calloc(nbins + 1, sizeof(int));
    cl->bin_atoms = (int *)malloc(sizeof(int) * (na > 0 ? na : 1));
    for (int i = 0; i < na; ++i) {
        double f[3];
        int b[3];
        for (int k = 0; k < 3; ++k)
            f[k] = pos[3 * i] * cl->inv[k] + pos[3 * i + 1] * cl->inv[3 + k] + pos[3 * i + 2] * cl->inv[6 + k];
        for (int k = 0; k < 3; ++k) {
            double fl = floor(f[k]);
            cl->shift[3 * i + k] = (int)fl;
            double fw = f[k] - fl;
            b[k] = (int)floor(fw * cl->nb[k]);
            if (b[k] >= cl->nb[k]) b[k] = cl->nb[k] - 1;
            if (b[k] < 0) b[k] = 0;
        }
        for (int c = 0; c < 3; ++c)
            cl->wpos[3 * i + c] = pos[3 * i + c] - cl->shift[3 * i] * lat[c] -
                                  cl->shift[3 * i + 1] * lat[3 + c] - cl->shift[3 * i + 2] * lat[6 + c];
        bin[i] = (b[0] * cl->nb[1] + b[1]) * cl->nb[2] + b[2];
        cl->bin_start[bin[i] + 1]++;
    }
    for (int t = 0; t < nbins; ++t) cl->bin_start[t + 1] += cl->bin_start[t];
    int *cur = (int *)malloc(sizeof(int) * nbins);
    memcpy(cur, cl->bin_start, sizeof(int) * nbins);
    for (int i = 0; i < na; ++i) cl->bin_atoms[cur[bin[i]]++] = i; /* ascending index per bin */
    free(cur);
    free(bin);
}

static void cell_free(cell_list *cl) {
    free(cl->wpos);
    free(cl->shift);
    free(cl->bin_start);
    free(cl->bin_atoms);
}

static int floordiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

/* All (j, image) within rcut of atom i; grows *buf as needed. */
static int cell_neighbors(const cell_list *cl, int i, double rcut, const double *w, nbr **buf,
                          int *cap) {
    int count = 0;
    double f[3];
    int b[3];
    for (int k = 0; k < 3; ++k) {
        f[k] = cl->wpos[3 * i] * cl->inv[k] + cl->wpos[3 * i + 1] * cl->inv[3 + k] +
               cl->wpos[3 * i + 2] * cl->inv[6 + k];
        b[k] = (int)floor(f[k] * cl->nb[k]);
        if (b[k] >= cl->nb[k]) b[k] = cl->nb[k] - 1;
        if (b[k] < 0) b[k] = 0;
    }
    double rc2 = rcut * rcut;
    for (int o0 = -cl->S[0]; o0 <= cl->S[0]; ++o0)
        for (int o1 = -cl->S[1]; o1 <= cl->S[1]; ++o1)
            for (int o2 = -cl->S[2]; o2 <= cl->S[2]; ++o2) {
                int bb[3] = {b[0] + o0, b[1] + o1, b[2] + o2}, t[3], wb[3];
                for (int k = 0; k < 3; ++k) {
                    t[k] = floordiv(bb[k], cl->nb[k]);
                    wb[k] = bb[k] - t[k] * cl->nb[k];
                }
                int bin = (wb[0] * cl->nb[1] + wb[1]) * cl->nb[2] + wb[2];
                double sh[3];
                for (int c = 0; c < 3; ++c) sh[c] = t[0] * cl->L[c] + t[1] * cl->L[3 + c] + t[2] * cl->L[6 + c];
                for (int q = cl->bin_start[bin]; q < cl->bin_start[bin + 1]; ++q) {
                    int j = cl->bin_atoms[q];
                    if (j == i && t[0] == 0 && t[1] == 0 && t[2] == 0) continue;
                    double d[3];
                    for (int c = 0; c < 3; ++c) d[c] = cl->wpos[3 * j + c] + sh[c] - cl->wpos[3 * i + c];
                    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                    if (r2 > rc2 * (1.0 + 1e-12)) continue;
                    double r = sqrt(r2);
                    if (r > rcut) continue;
                    if (count >= *cap) {
                        *cap = *cap * 2 + 64;
                        *buf = (nbr *)realloc(*buf, sizeof(nbr) * (*cap));
                    }
                    nbr *e = *buf + count++;
                    e->j = j;
                    for (int k = 0; k < 3; ++k)
                        e->n[k] = t[k] - cl->shift[3 * j + k] + cl->shift[3 * i + k];
                    e->d[0] = d[0];
                    e->d[1] = d[1];
                    e->d[2] = d[2];
                    e->r = r;
                    e->w = w ? w[j] : 0.0;
                }
            }
    return count;
}

/* ------------------------------------------------------------------------ */
/* sparse mode: per-atom fused forward / GPR / backward                     */
/* ------------------------------------------------------------------------ */

static void atom_features(const orc_model *m, const nbr *nb, int K, double *x) {
    const int nsf = m->nsf;
    for (int d = 0; d < 2 * nsf; ++d) x[d] = 0.0;
    for (int s = 0; s < nsf; ++s) {
        const int type = m->type[s];
        const double rc = m->cut[s];
        if (type == 1 || type == 3) {
            const double alpha = (type == 1) ? m->alpha[s] : 4.0;
            const double rs = (type == 1) ? 0.0 : m->alpha[s];
            for (int a = 0; a < K; ++a) {
                double r = nb[a].r;
                if (r > rc) continue;
                double term = exp(-alpha * (r - rs) * (r - rs)) * fcut(r, rc);
                x[s] += term;
                x[s + nsf] += term * nb[a].w;
            }
        } else if (type == 2 || type == 4) {
            const double eta = m->alpha[s], lam = (type == 2) ? 1.0 : -1.0, zeta = zeta_of(m, s);
            for (int a = 0; a < K; ++a) {
                double rij = nb[a].r;
                if (rij > rc) continue;
                double fij = fcut(rij, rc);
                for (int b = a + 1; b < K; ++b) {
                    double rik = nb[b].r;
                    if (rik > rc) continue;
                    double e0 = nb[b].d[0] - nb[a].d[0], e1 = nb[b].d[1] - nb[a].d[1], e2 = nb[b].d[2] - nb[a].d[2];
                    double rjk = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
                    if (rjk > rc) continue;
                    double cth = (rjk * rjk - rij * rij - rik * rik) / (-2.0 * rij * rik);
                    double P, dP;
                    ang_power(1.0 + lam * cth, lam, zeta, &P, &dP);
                    double term = P * exp(-eta * (rij * rij + rik * rik + rjk * rjk)) * fij * fcut(rik, rc) *
                                  fcut(rjk, rc);
                    x[s] += term;
                    x[s + nsf] += term * nb[a].w * nb[b].w;
                }
            }
        }
    }
}

/* Per-atom GPR in the reference's direct form; returns E_i, writes g = dE_i/dx. */
static double atom_gpr(const orc_model *m, const double *x, double *g) {
    const int D = m->des_len, M = m->nsparse;
    double e = 0.0;
    for (int d = 0; d < D; ++d) g[d] = 0.0;
    for (int k = 0; k < M; ++k) {
        const double *mk = m->mm + (size_t)k * D;
        double t = 0.0;
        for (int d = 0; d < D; ++d) {
            double u = (x[d] - mk[d]) / m->theta[d];
            t += u * u;
        }
        double kc = ORC_DELTA * exp(-0.5 * t) * m->coeff[k];
        e += kc;
        for (int d = 0; d < D; ++d) g[d] -= (x[d] - mk[d]) / (m->theta[d] * m->theta[d]) * kc;
    }
    return e;
}

/* Backward for centre atom i: fs[3a+c] = dE_i/dR(neighbour a), vir = sum T d d^T / r. */
static void atom_backward(const orc_model *m, const nbr *nb, int K, const double *g, double *fs,
                          double *vir) {
    const int nsf = m->nsf;
    double *T = (double *)calloc(K > 0 ? K : 1, sizeof(double));
    for (int a = 0; a < 3 * K; ++a) fs[a] = 0.0;
    for (int t = 0; t < 9; ++t) vir[t] = 0.0;
    for (int s = 0; s < nsf; ++s) {
        const int type = m->type[s];
        const double rc = m->cut[s];
        const double gs = g[s], gw = g[s + nsf];
        if (type == 1 || type == 3) {
            const double alpha = (type == 1) ? m->alpha[s] : 4.0;
            const double rs = (type == 1) ? 0.0 : m->alpha[s];
            for (int a = 0; a < K; ++a) {
                double r = nb[a].r;
                if (r > rc) continue;
                double ex = exp(-alpha * (r - rs) * (r - rs));
                double dterm = ex * (-2.0 * alpha * (r - rs) * fcut(r, rc) + dfcut(r, rc));
                T[a] += (gs + nb[a].w * gw) * dterm;
            }
        } else if (type == 2 || type == 4) {
            const double eta = m->alpha[s], lam = (type == 2) ? 1.0 : -1.0, zeta = zeta_of(m, s);
            for (int a = 0; a < K; ++a) {
                double ra = nb[a].r;
                if (ra > rc) continue;
                double fa = fcut(ra, rc), dfa = dfcut(ra, rc);
                for (int b = a + 1; b < K; ++b) {
                    double rb = nb[b].r;
                    if (rb > rc) continue;
                    double dj[3] = {nb[b].d[0] - nb[a].d[0], nb[b].d[1] - nb[a].d[1], nb[b].d[2] - nb[a].d[2]};
                    double re = sqrt(dj[0] * dj[0] + dj[1] * dj[1] + dj[2] * dj[2]);
                    if (re > rc) continue;
                    double fb = fcut(rb, rc), dfb = dfcut(rb, rc), fe = fcut(re, rc), dfe = dfcut(re, rc);
                    double cth = (ra * ra + rb * rb - re * re) / (2.0 * ra * rb);
                    double P, dP;
                    ang_power(1.0 + lam * cth, lam, zeta, &P, &dP);
                    double ex = exp(-eta * (ra * ra + rb * rb + re * re));
                    double F3 = fa * fb * fe;
                    double geff = gs + nb[a].w * nb[b].w * gw;
                    /* d cos / d(ra, rb, re) */
                    double dca = 1.0 / rb - cth / ra, dcb = 1.0 / ra - cth / rb, dce = -re / (ra * rb);
                    double Ta = geff * (dP * dca * ex * F3 - 2.0 * eta * ra * P * ex * F3 + P * ex * dfa * fb * fe);
                    double Tb = geff * (dP * dcb * ex * F3 - 2.0 * eta * rb * P * ex * F3 + P * ex * fa * dfb * fe);
                    double Te = geff * (dP * dce * ex * F3 - 2.0 * eta * re * P * ex * F3 + P * ex * fa * fb * dfe);
                    T[a] += Ta;
                    T[b] += Tb;
                    /* r_jk term: d re / d R_k = dj/re, d re / d R_j = -dj/re */
                    for (int c = 0; c < 3; ++c) {
                        fs[3 * b + c] += Te * dj[c] / re;
                        fs[3 * a + c] -= Te * dj[c] / re;
                    }
                    for (int p = 0; p < 3; ++p)
                        for (int q = 0; q < 3; ++q) vir[3 * p + q] += Te * dj[p] * dj[q] / re;
                }
            }
        }
    }
    for (int a = 0; a < K; ++a) {
        double r = nb[a].r;
        for (int c = 0; c < 3; ++c) fs[3 * a + c] += T[a] * nb[a].d[c] / r;
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) vir[3 * p + q] += T[a] * nb[a].d[p] * nb[a].d[q] / r;
    }
    free(T);
}

static int sparse_core(const orc_model *m, int na, const int *species, const double *lat,
                       const double *pos, double rcut, int lgrad, int nthreads, int i0, int i1,
                       double *ene, double *force, double *stress6, double *xx_out,
                       double *dedg_out, double *eatom_out) {
    if (!model_ok(m)) return -3;
    const int D = m->des_len;
    double *w = (double *)malloc(sizeof(double) * (na > 0 ? na : 1));
    if (orc_species_weights(m, na, species, w)) {
        free(w);
        return -2;
    }
    cell_list cl;
    cell_build(&cl, na, lat, pos, rcut);
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    int nt = 1;
    (void)nthreads;
#endif
    double *fpriv = (double *)calloc((size_t)nt * na * 3, sizeof(double));
    double *vpriv = (double *)calloc((size_t)nt * 9, sizeof(double));
    double *epriv = (double *)calloc((size_t)nt, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        int cap = 256;
        nbr *buf = (nbr *)malloc(sizeof(nbr) * cap);
        double *x = (double *)malloc(sizeof(double) * D);
        double *g = (double *)malloc(sizeof(double) * D);
        double *fs = NULL;
        int fcap = 0;
        double vir[9];
        double *fl = fpriv + (size_t)tid * na * 3;
#pragma omp for schedule(dynamic, 4)
        for (int i = i0; i < i1; ++i) {
            int K = cell_neighbors(&cl, i, rcut, w, &buf, &cap);
            atom_features(m, buf, K, x);
            double e = atom_gpr(m, x, g);
            epriv[tid] += e;
            if (xx_out) memcpy(xx_out + (size_t)i * D, x, sizeof(double) * D);
            if (dedg_out) memcpy(dedg_out + (size_t)i * D, g, sizeof(double) * D);
            if (eatom_out) eatom_out[i] = e;
            if (!lgrad) continue;
            if (3 * K > fcap) {
                fcap = 3 * K + 64;
                fs = (double *)realloc(fs, sizeof(double) * fcap);
            }
            atom_backward(m, buf, K, g, fs, vir);
            for (int a = 0; a < K; ++a)
                for (int c = 0; c < 3; ++c) {
                    fl[3 * buf[a].j + c] -= fs[3 * a + c]; /* F = -dE/dR */
                    fl[3 * i + c] += fs[3 * a + c];
                }
            for (int t = 0; t < 9; ++t) vpriv[(size_t)tid * 9 + t] += vir[t];
        }
        free(buf);
        free(x);
        free(g);
        free(fs);
    }
    double etot = 0.0, v[9] = {0};
    for (int a = 0; a < na * 3; ++a) force[a] = 0.0;
    for (int t = 0; t < nt; ++t) {
        etot += epriv[t];
        for (int a = 0; a < na * 3; ++a) force[a] += fpriv[(size_t)t * na * 3 + a];
        for (int q = 0; q < 9; ++q) v[q] += vpriv[(size_t)t * 9 + q];
    }
    double vol = fabs(det_lat(lat));
    double s = -(1.0 / ORC_GPA) / vol; /* stress_cal = -virial, gap_calc.f90:189-203 */
    stress6[0] = v[0] * s;
    stress6[1] = v[4] * s;
    stress6[2] = v[8] * s;
    stress6[3] = v[1] * s;
    stress6[4] = v[5] * s;
    stress6[5] = v[2] * s;
    *ene = etot;
    free(fpriv);
    free(vpriv);
    free(epriv);
    cell_free(&cl);
    free(w);
    return 0;
}

int orc_gap_calc_sparse(const orc_model *m, int na, const int *species, const double *lat,
                        const double *pos, double rcut, int lgrad, int nthreads, double *ene,
                        double *force, double *stress6, double *xx, double *dedg, double *eatom) {
    return sparse_core(m, na, species, lat, pos, rcut, lgrad, nthreads, 0, na, ene, force, stress6,
                       xx, dedg, eatom);
}

int orc_gap_calc_sparse_range(const orc_model *m, int na, const int *species,
                              const double *lat, const double *pos, double rcut, int lgrad,
                              int nthreads, int i0, int i1, double *ene, double *force,
                              double *stress6) {
    if (i0 < 0 || i1 > na || i0 > i1) return -4;
    return sparse_core(m, na, species, lat, pos, rcut, lgrad, nthreads, i0, i1, ene, force, stress6,
                       NULL, NULL, NULL);
}

int orc_descriptors(const orc_model *m, int na, const int *species, const double *lat,
                    const double *pos, double rcut, int nthreads, double *xx) {
    if (!model_ok(m)) return -3;
    const int D = m->des_len;
    double *w = (double *)malloc(sizeof(double) * (na > 0 ? na : 1));
    if (orc_species_weights(m, na, species, w)) {
        free(w);
        return -2;
    }
    cell_list cl;
    cell_build(&cl, na, lat, pos, rcut);
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    int nt = 1;
    (void)nthreads;
#endif
#pragma omp parallel num_threads(nt)
    {
        int cap = 256;
        nbr *buf = (nbr *)malloc(sizeof(nbr) * cap);
#pragma omp for schedule(dynamic, 4)
        for (int i = 0; i < na; ++i) {
            int K = cell_neighbors(&cl, i, rcut, w, &buf, &cap);
            atom_features(m, buf, K, xx + (size_t)i * D);
        }
        free(buf);
    }
    cell_free(&cl);
    free(w);
    return 0;
}

static int cmp_nb4(const void *a, const void *b) {
    const int *x = (const int *)a, *y = (const int *)b;
    for (int t = 0; t < 4; ++t)
        if (x[t] != y[t]) return x[t] < y[t] ? -1 : 1;
    return 0;
}

long orc_neighbor_set(int na, const double *lat, const double *pos, double rcut, long cap,
                      int *out4, double *dist, int *count_per_atom) {
    cell_list cl;
    cell_build(&cl, na, lat, pos, rcut);
    int bcap = 256;
    nbr *buf = (nbr *)malloc(sizeof(nbr) * bcap);
    long total = 0;
    int *tmp = NULL;
    int tcap = 0;
    for (int i = 0; i < na; ++i) {
        int K = cell_neighbors(&cl, i, rcut, NULL, &buf, &bcap);
        if (count_per_atom) count_per_atom[i] = K;
        if (5 * K > tcap) {
            tcap = 5 * K + 64;
            tmp = (int *)realloc(tmp, sizeof(int) * tcap);
        }
        /* sort by (j, n1, n2, n3); keep the index to carry the distance */
        for (int a = 0; a < K; ++a) {
            tmp[5 * a] = buf[a].j;
            tmp[5 * a + 1] = buf[a].n[0];
            tmp[5 * a + 2] = buf[a].n[1];
            tmp[5 * a + 3] = buf[a].n[2];
            tmp[5 * a + 4] = a;
        }
        qsort(tmp, K, sizeof(int) * 5, cmp_nb4);
        for (int a = 0; a < K; ++a, ++total) {
            if (total < cap) {
                for (int t = 0; t < 4; ++t) out4[4 * total + t] = tmp[5 * a + t];
                if (dist) dist[total] = buf[tmp[5 * a + 4]].r;
            }
        }
    }
    free(tmp);
    free(buf);
    cell_free(&cl);
    return total;
}
