/*
 * pbte_oracle.c -- CPU restatement of the reference's hot-path kernels, in
 * binary64.  TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and
 * the CPU-baseline leg of the benchmark link it as the checker.  Nothing in
 * the product path calls it.
 *
 * Every function follows one reference function of
 * solver-python/pybte/_kernels.py (cited per function) with the same loop
 * nest and the same expression order, so that in FP64 its sweeps reproduce
 * the reference's iterate sequence.  Parallelism is OpenMP over discrete
 * directions, the reference's only parallel loop (_kernels.py:24-30).
 *
 * Pinning (what checks each function, beyond retyping):
 *   lu_factor / lu_solve   scipy LAPACK getrf/getrs, bitwise (tests)
 *   sweep + moments + residual + wall_flux
 *                          the reference's golden converged vdf and residual /
 *                          mass histories (F1/F2/F3_1_1) and FROZEN counts
 *   transport_residual     zero at the certified fixed points; equals the
 *                          CIS step residual identity (sweep.py:169-180)
 *   order_scan             brute-force Fortran scan, checked against the
 *                          pass-number closed form and topological validity
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* Problem description shared by every entry point (all arrays C-contiguous,
 * index orders as the reference's SweepContext, sweep.py:37-72). */
typedef struct {
    int32_t ndir, n_tris, nd, n_faces;
    const double *cxv, *cyv, *domega;         /* [D]                         */
    double cv, vg, tau_r, tau_n, tau_c, pi;
    const double *mass;                       /* [P][P][N]  int_tri_tri      */
    const double *gx, *gy;                    /* [N][P][P]                   */
    const double *fcm, *ttfc;                 /* [N][3][P][P]                */
    const double *nx, *ny;                    /* [N][3]                      */
    const int32_t *tri_faces, *face_bc;       /* [N][3], [F]                 */
    const int32_t *neighbour;                 /* [N][3]                      */
    const int32_t *periodic_tri;              /* [F]                         */
    const int32_t *bc_type;                   /* [nbc]                       */
    const double *bc_temp;                    /* [nbc]                       */
    const double *flux_wall;                  /* [F][P]                      */
    const double *int_tri;                    /* [P][N]                      */
    const int32_t *face_tri, *face_lfc;       /* [F][2]                      */
} orc_problem;

/* DGETF2 on a row-major n x n block (_kernels.py:43-76). */
ORC_API int orc_lu_factor(int n, double *a, int32_t *ipiv)
{
    int info = 0;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double big = fabs(a[k * n + k]);
        for (int i = k + 1; i < n; ++i) {
            double v = fabs(a[i * n + k]);
            if (v > big) { big = v; p = i; }
        }
        ipiv[k] = p;
        if (a[p * n + k] == 0.0) {
            if (info == 0) info = k + 1;
            continue;
        }
        if (p != k)
            for (int j = 0; j < n; ++j) {
                double t = a[k * n + j]; a[k * n + j] = a[p * n + j]; a[p * n + j] = t;
            }
        double r = 1.0 / a[k * n + k];
        for (int i = k + 1; i < n; ++i) a[i * n + k] = a[i * n + k] * r;
        for (int i = k + 1; i < n; ++i) {
            double aik = a[i * n + k];
            if (aik != 0.0)
                for (int j = k + 1; j < n; ++j) a[i * n + j] -= aik * a[k * n + j];
        }
    }
    return info;
}

/* DGETRS('N') for one right-hand side (_kernels.py:78-101). */
ORC_API void orc_lu_solve(int n, const double *a, const int32_t *ipiv, double *b)
{
    for (int k = 0; k < n; ++k) {
        int p = ipiv[k];
        if (p != k) { double t = b[k]; b[k] = b[p]; b[p] = t; }
    }
    for (int k = 0; k < n; ++k) {
        double bk = b[k];
        if (bk != 0.0)
            for (int i = k + 1; i < n; ++i) b[i] -= bk * a[i * n + k];
    }
    for (int k = n - 1; k >= 0; --k) {
        b[k] = b[k] / a[k * n + k];
        double bk = b[k];
        if (bk != 0.0)
            for (int i = 0; i < k; ++i) b[i] -= bk * a[i * n + k];
    }
}

/* A_SOL of element i for one direction (_kernels.py:106-125). */
static void fill_a_sol(const orc_problem *p, double *a, int i, double cxd, double cyd,
                       double inv_tau_c)
{
    const int nd = p->nd, N = p->n_tris;
    double s0 = cxd * p->nx[i * 3 + 0] + cyd * p->ny[i * 3 + 0];
    double s1 = cxd * p->nx[i * 3 + 1] + cyd * p->ny[i * 3 + 1];
    double s2 = cxd * p->nx[i * 3 + 2] + cyd * p->ny[i * 3 + 2];
    double w0 = 0.5 * (s0 + fabs(s0)), w1 = 0.5 * (s1 + fabs(s1)), w2 = 0.5 * (s2 + fabs(s2));
    const double *fc = p->fcm + (size_t)i * 3 * nd * nd;
    for (int l = 0; l < nd; ++l)
        for (int m = 0; m < nd; ++m) {
            double v = (inv_tau_c * p->mass[((size_t)m * nd + l) * N + i]
                        - cxd * p->gx[((size_t)i * nd + m) * nd + l]
                        - cyd * p->gy[((size_t)i * nd + m) * nd + l]);
            v = v + w0 * fc[(0 * nd + m) * nd + l];
            v = v + w1 * fc[(1 * nd + m) * nd + l];
            v = v + w2 * fc[(2 * nd + m) * nd + l];
            a[m * nd + l] = v;
        }
}

/* Callaway scattering source, reference expression order (_kernels.py:153-167). */
static void fill_a_src(const orc_problem *p, double *src, int i, double cxd, double cyd,
                       const double *ts, const double *qxs, const double *qys)
{
    const int nd = p->nd, N = p->n_tris;
    const double cv = p->cv, vg = p->vg, pi = p->pi;
    for (int m = 0; m < nd; ++m) {
        double acc = 0.0;
        for (int l = 0; l < nd; ++l) {
            double mlm = p->mass[((size_t)l * nd + m) * N + i];
            double t = ts[(size_t)l * N + i];
            acc = acc + cv * t / 4.0 / pi / p->tau_r * mlm;
            acc = acc + 1.0 / p->tau_n * mlm *
                  (cv * t / 4.0 / pi +
                   3.0 / 4.0 / pi * (qxs[(size_t)l * N + i] * cxd + qys[(size_t)l * N + i] * cyd) / vg / vg);
        }
        src[m] = acc;
    }
}

/* Upwind inflow by boundary type (_kernels.py:169-215). */
static void add_face_sources(const orc_problem *p, double *src, int i, int d, double cxd,
                             double cyd, const double *vdf, const double *pbuf, int use_pbuf)
{
    const int nd = p->nd, N = p->n_tris;
    for (int il = 0; il < 3; ++il) {
        double speed = cxd * p->nx[i * 3 + il] + cyd * p->ny[i * 3 + il];
        double w = 0.5 * (speed - fabs(speed));
        if (w == 0.0) continue;
        int fc = p->tri_faces[i * 3 + il];
        int bc = p->face_bc[fc];
        const double *tt = p->ttfc + ((size_t)i * 3 + il) * nd * nd;
        if (bc < 0) {
            int iext = p->neighbour[i * 3 + il];
            const double *fe = vdf + ((size_t)d * N + iext) * nd;
            for (int l = 0; l < nd; ++l) {
                double fl = fe[l];
                for (int m = 0; m < nd; ++m) src[m] -= (w * tt[m * nd + l]) * fl;
            }
        } else {
            int typ = p->bc_type[bc];
            if (typ == 3) {
                int iext = p->periodic_tri[fc];
                for (int l = 0; l < nd; ++l) {
                    double fl = use_pbuf ? pbuf[((size_t)d * p->n_faces + fc) * nd + l]
                                         : vdf[((size_t)d * N + iext) * nd + l];
                    for (int m = 0; m < nd; ++m) src[m] -= (w * tt[m * nd + l]) * fl;
                }
            } else if (typ == 2) {
                for (int m = 0; m < nd; ++m) src[m] -= w * (-p->flux_wall[(size_t)fc * nd + m]);
            } else {
                double tw = p->cv / 4.0 / p->pi * p->bc_temp[bc];
                for (int m = 0; m < nd; ++m) src[m] -= w * (tw * tt[m * nd + 0]);
            }
        }
    }
}

/* Factorise A_SOL for every (direction, element) (_kernels.py:127-148).
 * lu: [D][N][P][P], piv: [D][N][P].  Returns the number of singular blocks. */
ORC_API int orc_build_operators(const orc_problem *p, double *lu, int32_t *piv)
{
    const int nd = p->nd, N = p->n_tris;
    const double inv_tau_c = 1.0 / p->tau_c;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
    for (int d = 0; d < p->ndir; ++d) {
        for (int i = 0; i < N; ++i) {
            double *a = lu + ((size_t)d * N + i) * nd * nd;
            fill_a_sol(p, a, i, p->cxv[d], p->cyv[d], inv_tau_c);
            if (orc_lu_factor(nd, a, piv + ((size_t)d * N + i) * nd) != 0) bad += 1;
        }
    }
    return bad;
}

/* One transport sweep with precomputed LU, in place (_kernels.py:217-239). */
ORC_API void orc_sweep(const orc_problem *p, const int32_t *order, const double *ts,
                       const double *qxs, const double *qys, const double *lu,
                       const int32_t *piv, double *vdf, const double *pbuf, int use_pbuf)
{
    const int nd = p->nd, N = p->n_tris;
#pragma omp parallel for schedule(dynamic, 1)
    for (int d = 0; d < p->ndir; ++d) {
        double src[16];
        const double cxd = p->cxv[d], cyd = p->cyv[d];
        for (int k = 0; k < N; ++k) {
            int i = order[(size_t)d * N + k];
            fill_a_src(p, src, i, cxd, cyd, ts, qxs, qys);
            add_face_sources(p, src, i, d, cxd, cyd, vdf, pbuf, use_pbuf);
            orc_lu_solve(nd, lu + ((size_t)d * N + i) * nd * nd, piv + ((size_t)d * N + i) * nd, src);
            for (int m = 0; m < nd; ++m) vdf[((size_t)d * N + i) * nd + m] = src[m];
        }
    }
}

/* ||A_SOL f - A_SRC||^2 per direction at the current state (_kernels.py:274-306). */
ORC_API void orc_transport_residual(const orc_problem *p, const double *ts, const double *qxs,
                                    const double *qys, const double *vdf, double *out_num,
                                    double *out_den)
{
    const int nd = p->nd, N = p->n_tris;
    const double inv_tau_c = 1.0 / p->tau_c;
#pragma omp parallel for schedule(dynamic, 1)
    for (int d = 0; d < p->ndir; ++d) {
        double src[16], a[256];
        double num = 0.0, den = 0.0;
        const double cxd = p->cxv[d], cyd = p->cyv[d];
        for (int i = 0; i < N; ++i) {
            fill_a_sol(p, a, i, cxd, cyd, inv_tau_c);
            fill_a_src(p, src, i, cxd, cyd, ts, qxs, qys);
            add_face_sources(p, src, i, d, cxd, cyd, vdf, NULL, 0);
            const double *f = vdf + ((size_t)d * N + i) * nd;
            for (int m = 0; m < nd; ++m) {
                double r = -src[m];
                for (int l = 0; l < nd; ++l) r += a[m * nd + l] * f[l];
                num += r * r * p->domega[d];
                den += src[m] * src[m] * p->domega[d];
            }
        }
        out_num[d] = num;
        out_den[d] = den;
    }
}

/* Angular moments and element integrals (_kernels.py:311-338). */
ORC_API void orc_moments(const orc_problem *p, const double *vdf, double *ts, double *qxs,
                         double *qys, double *temp, double *qx, double *qy)
{
    const int nd = p->nd, N = p->n_tris, D = p->ndir;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        double t_i = 0.0, qx_i = 0.0, qy_i = 0.0;
        for (int l = 0; l < nd; ++l) {
            double ss = 0.0, ssx = 0.0, ssy = 0.0;
            for (int d = 0; d < D; ++d) {
                double f = vdf[((size_t)d * N + i) * nd + l];
                ss += f * p->domega[d];
                ssx += p->cxv[d] * f * p->domega[d];
                ssy += p->cyv[d] * f * p->domega[d];
            }
            ts[(size_t)l * N + i] = ss / p->cv;
            qxs[(size_t)l * N + i] = ssx;
            qys[(size_t)l * N + i] = ssy;
            double w = p->int_tri[(size_t)l * N + i];
            t_i += ss * w / p->cv;
            qx_i += ssx * w;
            qy_i += ssy * w;
        }
        temp[i] = t_i;
        qx[i] = qx_i;
        qy[i] = qy_i;
    }
}

/* Sequential sum and the iterate residual (_kernels.py:340-361). */
/* Threads of the OpenMP loops from now on (the CPU-baseline timing runs
 * the sweep at the host's thread count and at one thread). */
ORC_API void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

ORC_API double orc_sum(const double *a, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i];
    return s;
}

ORC_API double orc_residual(const double *temp, const double *t_old, int n)
{
    double s1 = 0.0, s2 = 0.0;
    for (int i = 0; i < n; ++i) {
        double d = temp[i] - t_old[i];
        s1 += d * d;
        s2 += temp[i] * temp[i];
    }
    return sqrt(s1 / s2);
}

/* Diffuse adiabatic wall emission (_kernels.py:363-397). flux_wall [F][P]. */
ORC_API void orc_wall_flux(const orc_problem *p, const double *vdf, double *flux_wall)
{
    const int nd = p->nd, N = p->n_tris, D = p->ndir;
#pragma omp parallel for schedule(dynamic, 16)
    for (int f = 0; f < p->n_faces; ++f) {
        int bc = p->face_bc[f];
        if (bc < 0 || p->bc_type[bc] != 2) continue;
        int tri, lfc;
        if (p->face_tri[f * 2 + 0] < 0) { tri = p->face_tri[f * 2 + 1]; lfc = p->face_lfc[f * 2 + 1]; }
        else { tri = p->face_tri[f * 2 + 0]; lfc = p->face_lfc[f * 2 + 0]; }
        double n1 = -p->nx[tri * 3 + lfc], n2 = -p->ny[tri * 3 + lfc];
        const double *fc = p->fcm + ((size_t)tri * 3 + lfc) * nd * nd;
        for (int m = 0; m < nd; ++m) {
            double ss1 = 0.0, ss2 = 0.0;
            for (int d = 0; d < D; ++d) {
                double u = p->cxv[d] * n1 + p->cyv[d] * n2;
                if (u < 0.0) {
                    double acc = 0.0;
                    for (int l = 0; l < nd; ++l) acc += vdf[((size_t)d * N + tri) * nd + l] * fc[m * nd + l];
                    ss1 += u * p->domega[d] * acc;
                } else {
                    ss2 -= u * p->domega[d];
                }
            }
            flux_wall[(size_t)f * nd + m] = -ss1 / ss2;
        }
    }
}

/* The Fortran scan topological sort, literally (ordering.py:61-140): repeated
 * ascending scans emitting every element whose inflow flags are clear and
 * clearing the reciprocal flags at once; a pass that emits nothing either
 * fails (allow_break == 0) or force-emits the queued element with the fewest
 * remaining inflow faces.  order: [D][N]; cycles: [D].  Returns 0 or -1. */
ORC_API int orc_order_scan(int ndir, int n_tris, const int32_t *neighbour, const double *nx,
                           const double *ny, const double *cx, const double *cy,
                           int32_t *order, int32_t *cycles, int allow_break)
{
    int status = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int d = 0; d < ndir; ++d) {
        int8_t *flag = (int8_t *)calloc((size_t)n_tris * 3, 1);
        uint8_t *queued = (uint8_t *)malloc((size_t)n_tris);
        memset(queued, 1, (size_t)n_tris);
        for (int i = 0; i < n_tris; ++i)
            for (int il = 0; il < 3; ++il)
                if (neighbour[i * 3 + il] >= 0) {
                    double fl = cx[d] * nx[i * 3 + il] + cy[d] * ny[i * 3 + il];
                    flag[i * 3 + il] = fl > 0.0 ? 1 : (fl < 0.0 ? -1 : 0);
                }
        int num = 0, broken = 0, fail = 0;
        while (num < n_tris) {
            int progressed = 0;
            for (int i = 0; i < n_tris; ++i) {
                if (!queued[i]) continue;
                if (flag[i * 3] == -1 || flag[i * 3 + 1] == -1 || flag[i * 3 + 2] == -1) continue;
                order[(size_t)d * n_tris + num++] = i;
                queued[i] = 0;
                progressed = 1;
                for (int il = 0; il < 3; ++il) {
                    int t2 = neighbour[i * 3 + il];
                    if (t2 < 0) continue;
                    for (int k = 0; k < 3; ++k)
                        if (neighbour[t2 * 3 + k] == i) { flag[t2 * 3 + k] = 0; break; }
                }
            }
            if (progressed) continue;
            if (!allow_break) { fail = 1; break; }
            int best = -1, best_in = 4;
            for (int i = 0; i < n_tris; ++i) {
                if (!queued[i]) continue;
                int cnt = (flag[i * 3] == -1) + (flag[i * 3 + 1] == -1) + (flag[i * 3 + 2] == -1);
                if (cnt < best_in) { best_in = cnt; best = i; }
            }
            if (best < 0) { fail = 1; break; }
            order[(size_t)d * n_tris + num++] = best;
            queued[best] = 0;
            broken += 1;
            for (int il = 0; il < 3; ++il) {
                int t2 = neighbour[best * 3 + il];
                if (t2 < 0) continue;
                for (int k = 0; k < 3; ++k)
                    if (neighbour[t2 * 3 + k] == best) { flag[t2 * 3 + k] = 0; break; }
            }
        }
        cycles[d] = broken;
        if (fail) {
#pragma omp atomic write
            status = -1;
            for (int k = num; k < n_tris; ++k) order[(size_t)d * n_tris + k] = -1;
        }
        free(flag);
        free(queued);
    }
    return status;
}
