/*
 * fk_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * This file is a plain-C restatement of the reference's arithmetic for the
 * Fisher-Kolmogorov forward-solve + log-likelihood path.  It is the checker
 * that the HIP product is compared against; nothing in the product links,
 * loads or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may use it.
 *
 * Built with gcc -O2 -ffp-contract=off (no -march flags), so every double
 * and float operation rounds exactly as the reference's x86-64 SSE build.
 *
 * Pins (see tests/test_oracle_*.py):
 *   - fk_* against the reference's own Glioma_ReactionDiffusionOperator /
 *     UpdateTumor compiled from source in oracle/_ref (bit-exact, both
 *     Real=float and Real=double), plus closed forms (pure reaction Euler
 *     recurrence, discrete Fourier decay, flux telescoping, CSF quirk).
 *   - og_* (ghost expansion) against the reference's MRAG_BBInfoCreator
 *     compiled from source in oracle/_ref on toy hierarchies, plus
 *     polynomial exactness of the interpolating wavelets.
 *   - fwt_* against the mathematics (tests/test_oracle_math_pins.py):
 *     details vanish on polynomials of degree < order per variable, equal
 *     the Lagrange interpolation remainder on x^order and x^(order+1), and
 *     single spikes give the Ga taps; code 0 is excluded from the maximum.
 *   - dump_uniform_f64 against an independent covering-cell formulation
 *     (replication factors 1/2/4 and the holes of coarser leaves).
 *   - lik_ref_* against the reference's likelihood program compiled from
 *     source in oracle/_ref, plus closed forms.
 *   - lik_flip_* against closed forms (constant fields, brute force).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* a1: Glioma_ReactionDiffusionOperator<3>::operator(), Case2 branch          */
/*     Glioma/Glioma_ReactionDiffusionOperator.h:148-218                      */
/* a2: UpdateTumor<3>::operator()  Glioma_ReactionDiffusionOperator.h:254-261 */
/* ------------------------------------------------------------------------ */

/* FP64 variant (Real = double): every temporary is double. */
static inline double fk_cell_f64(const double *phi, const double *pw, const double *pg,
                                 long c, long sx, long sy, long sz,
                                 double ih2, double Dw, double Dg, double rho)
{
    const long W = c - sx, E = c + sx, S = c - sy, N = c + sy, Bk = c - sz, F = c + sz;
    double df[6] = {0, 0, 0, 0, 0, 0};
    const double tissue = pw[c] + pg[c] + phi[c];              /* :155 */
    if (tissue > 0.) {
        const double brainW = pw[W] + pg[W];                    /* :174-180 */
        const double brain0 = pw[c] + pg[c];
        const double brainE = pw[E] + pg[E];
        const double brainS = pw[S] + pg[S];
        const double brainN = pw[N] + pg[N];
        const double brainB = pw[Bk] + pg[Bk];
        const double brainF = pw[F] + pg[F];
        /* :183-188 (the centre term is re-evaluated per face, same value) */
        df[0] = 2.0 * (1.0 / ((1.0 / ((pw[c] * Dw + pg[c] * Dg) / brain0)) + (1.0 / ((pw[W] * Dw + pg[W] * Dg) / brainW))));
        df[1] = 2.0 * (1.0 / ((1.0 / ((pw[c] * Dw + pg[c] * Dg) / brain0)) + (1.0 / ((pw[E] * Dw + pg[E] * Dg) / brainE))));
        df[2] = 2.0 * (1.0 / ((1.0 / ((pw[c] * Dw + pg[c] * Dg) / brain0)) + (1.0 / ((pw[S] * Dw + pg[S] * Dg) / brainS))));
        df[3] = 2.0 * (1.0 / ((1.0 / ((pw[c] * Dw + pg[c] * Dg) / brain0)) + (1.0 / ((pw[N] * Dw + pg[N] * Dg) / brainN))));
        df[4] = 2.0 * (1.0 / ((1.0 / ((pw[c] * Dw + pg[c] * Dg) / brain0)) + (1.0 / ((pw[Bk] * Dw + pg[Bk] * Dg) / brainB))));
        df[5] = 2.0 * (1.0 / ((1.0 / ((pw[c] * Dw + pg[c] * Dg) / brain0)) + (1.0 / ((pw[F] * Dw + pg[F] * Dg) / brainF))));
    }
    /* :198-203 Neumann doubling, sequential */
    if (df[0] == 0 && phi[W] > 0.) df[1] *= 2.0;
    if (df[1] == 0 && phi[E] > 0.) df[0] *= 2.0;
    if (df[2] == 0 && phi[S] > 0.) df[3] *= 2.0;
    if (df[3] == 0 && phi[N] > 0.) df[2] *= 2.0;
    if (df[4] == 0 && phi[Bk] > 0.) df[5] *= 2.0;
    if (df[5] == 0 && phi[F] > 0.) df[4] *= 2.0;
    /* :206-216 */
    const double in = ih2 * (df[0] * phi[W] + df[1] * phi[E] + df[2] * phi[S] + df[3] * phi[N] +
                             df[4] * phi[Bk] + df[5] * phi[F]);
    const double out = -((df[0] + df[1] + df[2] + df[3] + df[4] + df[5]) * phi[c] * ih2);
    const double reac = rho * phi[c] * (1. - phi[c]);
    return out + in + reac;
}

/* Real = float variant: storage and the marked temporaries are float, the
 * rest is double, exactly as the reference's default build (_MRAG_REAL_ =
 * float, MRAGcore/MRAGCommon.h:43-48).  Dw, Dg, rho are `const Real`
 * members of the operator (:18). */
static inline float fk_cell_f32ref(const float *phi, const float *pw, const float *pg,
                                   long c, long sx, long sy, long sz,
                                   double ih2, float Dw, float Dg, float rho)
{
    const long W = c - sx, E = c + sx, S = c - sy, N = c + sy, Bk = c - sz, F = c + sz;
    double df[6] = {0, 0, 0, 0, 0, 0};
    const float tissue = pw[c] + pg[c] + phi[c];
    if (tissue > 0.) {
        const float brainW = pw[W] + pg[W];
        const float brain0 = pw[c] + pg[c];
        const float brainE = pw[E] + pg[E];
        const float brainS = pw[S] + pg[S];
        const float brainN = pw[N] + pg[N];
        const float brainB = pw[Bk] + pg[Bk];
        const float brainF = pw[F] + pg[F];
        const float k0 = (pw[c] * Dw + pg[c] * Dg) / brain0;
        const float kW = (pw[W] * Dw + pg[W] * Dg) / brainW;
        const float kE = (pw[E] * Dw + pg[E] * Dg) / brainE;
        const float kS = (pw[S] * Dw + pg[S] * Dg) / brainS;
        const float kN = (pw[N] * Dw + pg[N] * Dg) / brainN;
        const float kB = (pw[Bk] * Dw + pg[Bk] * Dg) / brainB;
        const float kF = (pw[F] * Dw + pg[F] * Dg) / brainF;
        df[0] = 2.0 * (1.0 / ((1.0 / (double)k0) + (1.0 / (double)kW)));
        df[1] = 2.0 * (1.0 / ((1.0 / (double)k0) + (1.0 / (double)kE)));
        df[2] = 2.0 * (1.0 / ((1.0 / (double)k0) + (1.0 / (double)kS)));
        df[3] = 2.0 * (1.0 / ((1.0 / (double)k0) + (1.0 / (double)kN)));
        df[4] = 2.0 * (1.0 / ((1.0 / (double)k0) + (1.0 / (double)kB)));
        df[5] = 2.0 * (1.0 / ((1.0 / (double)k0) + (1.0 / (double)kF)));
    }
    if (df[0] == 0 && phi[W] > 0.) df[1] *= 2.0;
    if (df[1] == 0 && phi[E] > 0.) df[0] *= 2.0;
    if (df[2] == 0 && phi[S] > 0.) df[3] *= 2.0;
    if (df[3] == 0 && phi[N] > 0.) df[2] *= 2.0;
    if (df[4] == 0 && phi[Bk] > 0.) df[5] *= 2.0;
    if (df[5] == 0 && phi[F] > 0.) df[4] *= 2.0;
    const double in = ih2 * (df[0] * (double)phi[W] + df[1] * (double)phi[E] + df[2] * (double)phi[S] +
                             df[3] * (double)phi[N] + df[4] * (double)phi[Bk] + df[5] * (double)phi[F]);
    const double out = -((df[0] + df[1] + df[2] + df[3] + df[4] + df[5]) * (double)phi[c] * ih2);
    const float rp = rho * phi[c];
    const double reac = (double)rp * (1. - (double)phi[c]);
    return (float)(out + in + reac);
}

/* std::max((Real)0, x) then std::min((Real)1, x): NaN -> 0. */
static inline double clamp01_f64(double x)
{
    x = (0. < x) ? x : 0.;
    x = (x < 1.) ? x : 1.;
    return x;
}
static inline float clamp01_f32(float x)
{
    x = (0.f < x) ? x : 0.f;
    x = (x < 1.f) ? x : 1.f;
    return x;
}

/* One block on a ghosted lab with a 1-cell halo: lab arrays are
 * (B+2)^3, x fastest; dphidt is B^3.  h = info.h[0]. */
void fk_block_f64(const double *lab_phi, const double *lab_pw, const double *lab_pg, int B,
                  double h, double Dw, double Dg, double rho, double *dphidt)
{
    const long L = B + 2, sx = 1, sy = L, sz = L * L;
    const double ih2 = 1. / (h * h); /* :40-41 */
    for (int iz = 0; iz < B; iz++)
        for (int iy = 0; iy < B; iy++)
            for (int ix = 0; ix < B; ix++) {
                const long c = (ix + 1) + sy * (iy + 1) + sz * (iz + 1);
                dphidt[ix + B * (iy + B * iz)] = fk_cell_f64(lab_phi, lab_pw, lab_pg, c, sx, sy, sz, ih2, Dw, Dg, rho);
            }
}

void fk_block_f32ref(const float *lab_phi, const float *lab_pw, const float *lab_pg, int B,
                     float h, float Dw, float Dg, float rho, float *dphidt)
{
    const long L = B + 2, sx = 1, sy = L, sz = L * L;
    const double hd = h;
    const double ih2 = 1. / (hd * hd);
    for (int iz = 0; iz < B; iz++)
        for (int iy = 0; iy < B; iy++)
            for (int ix = 0; ix < B; ix++) {
                const long c = (ix + 1) + sy * (iy + 1) + sz * (iz + 1);
                dphidt[ix + B * (iy + B * iz)] = fk_cell_f32ref(lab_phi, lab_pw, lab_pg, c, sx, sy, sz, ih2, Dw, Dg, rho);
            }
}

void fk_update_f64(double *phi, const double *dphidt, long n, double dt)
{
    for (long i = 0; i < n; i++) phi[i] = clamp01_f64(phi[i] + dt * dphidt[i]);
}

void fk_update_f32ref(float *phi, const float *dphidt, long n, double dt)
{
    for (long i = 0; i < n; i++) {
        float p = (float)((double)phi[i] + dt * (double)dphidt[i]);
        phi[i] = clamp01_f32(p);
    }
}

/* ------------------------------------------------------------------------ */
/* Uniform periodic block grid (the 128^3 plain and adjoint workloads): N^3 cells as (N/B)^3 blocks */
/* of B^3.  Restates Glioma_HG_UQ::_reactionDiffusionStep (Glioma_HG_UQ.cpp:  */
/* 266-277): per block, BlockLab::load of a ghosted copy (same-level periodic */
/* neighbours, MRAGBlockLab.h:154-301), the RD functor, then a separate       */
/* UpdateTumor pass over all blocks.  Volumes are x-fastest N^3 arrays.       */
/* ------------------------------------------------------------------------ */
static inline long wrap(long i, long n) { i %= n; return i < 0 ? i + n : i; }

static void lab_gather_f64(const double *vol, long N, int B, long bx, long by, long bz, double *lab)
{
    const long L = B + 2;
    for (long z = -1; z <= B; z++)
        for (long y = -1; y <= B; y++)
            for (long x = -1; x <= B; x++) {
                const long gx = wrap(bx * B + x, N), gy = wrap(by * B + y, N), gz = wrap(bz * B + z, N);
                lab[(x + 1) + L * ((y + 1) + L * (z + 1))] = vol[gx + N * (gy + N * gz)];
            }
}

static void lab_gather_f32(const float *vol, long N, int B, long bx, long by, long bz, float *lab)
{
    const long L = B + 2;
    for (long z = -1; z <= B; z++)
        for (long y = -1; y <= B; y++)
            for (long x = -1; x <= B; x++) {
                const long gx = wrap(bx * B + x, N), gy = wrap(by * B + y, N), gz = wrap(bz * B + z, N);
                lab[(x + 1) + L * ((y + 1) + L * (z + 1))] = vol[gx + N * (gy + N * gz)];
            }
}

/* n_steps explicit Euler steps of the uniform grid, FP64. work: 3*(B+2)^3 per
 * thread is allocated internally; dphidt is an N^3 scratch. Returns 0. */
int fk_uniform_steps_f64(double *phi, const double *pw, const double *pg, long N, int B,
                         double h, double Dw, double Dg, double rho, double dt, int n_steps, int n_threads)
{
    const long nb = N / B;
    const long L = B + 2;
    double *dphidt = (double *)malloc(sizeof(double) * N * N * N);
    if (!dphidt) return -1;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
    for (int s = 0; s < n_steps; s++) {
#pragma omp parallel
        {
            double *lab = (double *)malloc(sizeof(double) * 3 * L * L * L);
            double *blk = (double *)malloc(sizeof(double) * B * B * B);
#pragma omp for schedule(dynamic, 4)
            for (long ib = 0; ib < nb * nb * nb; ib++) {
                const long bx = ib % nb, by = (ib / nb) % nb, bz = ib / (nb * nb);
                lab_gather_f64(phi, N, B, bx, by, bz, lab);
                lab_gather_f64(pw, N, B, bx, by, bz, lab + L * L * L);
                lab_gather_f64(pg, N, B, bx, by, bz, lab + 2 * L * L * L);
                fk_block_f64(lab, lab + L * L * L, lab + 2 * L * L * L, B, h, Dw, Dg, rho, blk);
                for (long z = 0; z < B; z++)
                    for (long y = 0; y < B; y++)
                        for (long x = 0; x < B; x++)
                            dphidt[(bx * B + x) + N * ((by * B + y) + N * (bz * B + z))] = blk[x + B * (y + B * z)];
            }
            free(lab);
            free(blk);
        }
        /* UpdateTumor: separate pass (BlockProcessing::process) */
#pragma omp parallel for schedule(static)
        for (long i = 0; i < N * N * N; i++) phi[i] = clamp01_f64(phi[i] + dt * dphidt[i]);
    }
    free(dphidt);
    return 0;
}

int fk_uniform_steps_f32ref(float *phi, const float *pw, const float *pg, long N, int B,
                            float h, float Dw, float Dg, float rho, double dt, int n_steps, int n_threads)
{
    const long nb = N / B;
    const long L = B + 2;
    float *dphidt = (float *)malloc(sizeof(float) * N * N * N);
    if (!dphidt) return -1;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
    for (int s = 0; s < n_steps; s++) {
#pragma omp parallel
        {
            float *lab = (float *)malloc(sizeof(float) * 3 * L * L * L);
            float *blk = (float *)malloc(sizeof(float) * B * B * B);
#pragma omp for schedule(dynamic, 4)
            for (long ib = 0; ib < nb * nb * nb; ib++) {
                const long bx = ib % nb, by = (ib / nb) % nb, bz = ib / (nb * nb);
                lab_gather_f32(phi, N, B, bx, by, bz, lab);
                lab_gather_f32(pw, N, B, bx, by, bz, lab + L * L * L);
                lab_gather_f32(pg, N, B, bx, by, bz, lab + 2 * L * L * L);
                fk_block_f32ref(lab, lab + L * L * L, lab + 2 * L * L * L, B, h, Dw, Dg, rho, blk);
                for (long z = 0; z < B; z++)
                    for (long y = 0; y < B; y++)
                        for (long x = 0; x < B; x++)
                            dphidt[(bx * B + x) + N * ((by * B + y) + N * (bz * B + z))] = blk[x + B * (y + B * z)];
            }
            free(lab);
            free(blk);
        }
#pragma omp parallel for schedule(static)
        for (long i = 0; i < N * N * N; i++) {
            float p = (float)((double)phi[i] + dt * (double)dphidt[i]);
            phi[i] = clamp01_f32(p);
        }
    }
    free(dphidt);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Adaptive block grid: point-value evaluation across levels.                */
/*                                                                           */
/* Restates the value that MRAG's ghost machinery assigns to a point        */
/* (MRAG_BBInfoCreator.inl:20-1315, MRAG_SmartBlockFinder.h:219-339):        */
/* a point p at level lam is located (vertex-centred, bIsCellCentered=false) */
/* in the leaf whose region contains its position; if that leaf has level   */
/* lam the stored value is used; if the leaf is finer, the value is the     */
/* level lam+1 point 2p (analysis with Ha = {1}); if coarser, it is the     */
/* synthesis sum_i Hs(2i - p) value(lam-1, i) (tensor product, zero weights */
/* skipped as in BastardGhost::generateRequests).  The domain is periodic.  */
/*                                                                           */
/* The expansion produces a list of (leaf, cell, weight) terms; terms that  */
/* hit the same leaf point are merged (weights are dyadic, so the merge is  */
/* exact) and the value is summed in a canonical order: by leaf level, then */
/* by global z, y, x of the point at that level.  The product's ghost       */
/* tables use the same canonical order, so FP64 ghost values agree bit for  */
/* bit between the two.                                                      */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n_leaves;
    int max_level;
    int B;
    const int *level;   /* [n_leaves] */
    const int *index;   /* [n_leaves*3] */
    const int *lut;     /* [(2^max_level)^3]: leaf covering each finest block cell */
    int order;          /* 2 or 4 */
} og_grid;

typedef struct {
    int leaf;
    int cell;
    double w;
    long key[4];
} og_term;

typedef struct {
    og_term *t;
    int n, cap;
    int overflow;
} og_terms;

static void og_push(og_terms *T, int leaf, int cell, double w, int lev, long gx, long gy, long gz)
{
    if (T->n == T->cap) { T->overflow = 1; return; }
    og_term *e = &T->t[T->n++];
    e->leaf = leaf; e->cell = cell; e->w = w;
    e->key[0] = lev; e->key[1] = gz; e->key[2] = gy; e->key[3] = gx;
}

/* Hs filters, MRAGWavelets_StaticData.cpp:16 and :34 (index = i - HsSupport[0]). */
static double og_hs(int order, int i)
{
    if (order == 2) {
        if (i < -1 || i > 1) return 0.0;
        static const double h2[3] = {0.5, 1., 0.5};
        return h2[i + 1];
    }
    if (i < -3 || i > 3) return 0.0;
    static const double h4[7] = {-1 / 16., 0, 9 / 16., 1, 9 / 16., 0, -1 / 16.};
    return h4[i + 3];
}

static int og_hs_lo(int order) { return order == 2 ? -1 : -3; }
static int og_hs_hi(int order) { return order == 2 ? 2 : 4; } /* exclusive */

static long floordiv(long a, long b) { long q = a / b; if ((a % b != 0) && ((a < 0) != (b < 0))) q--; return q; }

/* leaf containing level-lam point p (global index at level lam, unwrapped). */
static int og_lying_leaf(const og_grid *g, int lam, const long p[3])
{
    const long nF = 1L << g->max_level;
    long f[3];
    for (int d = 0; d < 3; d++) {
        /* position in finest-level block units: p * 2^(maxL-lam) / B */
        long q = p[d];
        if (g->max_level >= lam) q = q * (1L << (g->max_level - lam));
        else q = floordiv(q, 1L << (lam - g->max_level));
        f[d] = wrap(floordiv(q, g->B), nF);
    }
    return g->lut[f[0] + nF * (f[1] + nF * f[2])];
}

static void og_expand(const og_grid *g, int lam, const long p[3], double w, og_terms *T, int depth)
{
    if (depth > 64 || T->overflow) { T->overflow = 1; return; }
    const int leaf = og_lying_leaf(g, lam, p);
    const int ll = g->level[leaf];
    if (ll == lam) {
        const long n = (long)g->B << lam; /* points per dim at level lam */
        long q[3];
        int loc[3];
        for (int d = 0; d < 3; d++) {
            q[d] = wrap(p[d], n);
            loc[d] = (int)(q[d] - (long)g->index[3 * leaf + d] * g->B);
        }
        og_push(T, leaf, loc[0] + g->B * (loc[1] + g->B * loc[2]), w, lam, q[0], q[1], q[2]);
        return;
    }
    if (ll > lam) { /* finer leaf: analysis, Ha = {1} at support [0,1) */
        const long q[3] = {2 * p[0], 2 * p[1], 2 * p[2]};
        og_expand(g, lam + 1, q, w, T, depth + 1);
        return;
    }
    /* coarser leaf: synthesis from level lam-1 */
    long s[3], e[3];
    for (int d = 0; d < 3; d++) {
        s[d] = -floordiv(-(p[d] + og_hs_lo(g->order)), 2);         /* ceil((p+HsS0)/2) */
        e[d] = 1 + floordiv(p[d] + og_hs_hi(g->order) - 1, 2);      /* 1+floor((p+HsS1-1)/2) */
    }
    long i[3];
    for (i[2] = s[2]; i[2] < e[2]; i[2]++)
        for (i[1] = s[1]; i[1] < e[1]; i[1]++)
            for (i[0] = s[0]; i[0] < e[0]; i[0]++) {
                const double wx = og_hs(g->order, (int)(2 * i[0] - p[0]));
                const double wy = og_hs(g->order, (int)(2 * i[1] - p[1]));
                const double wz = og_hs(g->order, (int)(2 * i[2] - p[2]));
                const double ww = wx * wy * wz;
                if (ww == 0.0) continue;
                og_expand(g, lam - 1, i, w * ww, T, depth + 1);
            }
}

static int og_term_cmp(const void *a, const void *b)
{
    const og_term *x = (const og_term *)a, *y = (const og_term *)b;
    for (int k = 0; k < 4; k++) {
        if (x->key[k] < y->key[k]) return -1;
        if (x->key[k] > y->key[k]) return 1;
    }
    return 0;
}

/* Merge + canonical sort.  Returns number of merged terms (<= cap), or -1 on
 * overflow. */
static int og_canonical(og_terms *T)
{
    if (T->overflow) return -1;
    qsort(T->t, T->n, sizeof(og_term), og_term_cmp);
    int m = 0;
    for (int k = 0; k < T->n; k++) {
        if (m > 0 && og_term_cmp(&T->t[m - 1], &T->t[k]) == 0) T->t[m - 1].w += T->t[k].w;
        else T->t[m++] = T->t[k];
    }
    int r = 0;
    for (int k = 0; k < m; k++)
        if (T->t[k].w != 0.0) T->t[r++] = T->t[k];
    T->n = r;
    return r;
}

/* Public: expansion of one point of block `leaf` at local (possibly ghost)
 * coordinates loc[3] (level of the block).  Outputs merged canonical terms.
 * Returns the term count or -1 on overflow. */
int og_point_terms(int n_leaves, int max_level, int B, const int *level, const int *index, const int *lut,
                   int order, int leaf, const int loc[3], int cap, int *out_leaf, int *out_cell, double *out_w)
{
    og_grid g = {n_leaves, max_level, B, level, index, lut, order};
    og_term *buf = (og_term *)malloc(sizeof(og_term) * (size_t)(cap > 0 ? cap * 4 : 1));
    og_terms T = {buf, 0, cap * 4, 0};
    const int lam = level[leaf];
    const long p[3] = {(long)index[3 * leaf] * B + loc[0], (long)index[3 * leaf + 1] * B + loc[1],
                       (long)index[3 * leaf + 2] * B + loc[2]};
    og_expand(&g, lam, p, 1.0, &T, 0);
    int m = og_canonical(&T);
    if (m > cap) m = -1;
    for (int k = 0; k < m; k++) {
        out_leaf[k] = buf[k].leaf;
        out_cell[k] = buf[k].cell;
        out_w[k] = buf[k].w;
    }
    free(buf);
    return m;
}

/* Fill a ghosted lab (stencil [s0, s1) in every dim: size B+s1-s0-1) of one
 * field for one leaf.  A ghost value is the FP64 sum acc + w*v over the
 * expansion terms in depth-first order (synthesis children in increasing
 * z, y, x; duplicates not merged) -- the order the device kernels use, so
 * both produce the same bits.  data is [n_leaves][B^3] (double).  Returns 0,
 * or -1 on overflow. */
int og_fill_lab_f64(int n_leaves, int max_level, int B, const int *level, const int *index, const int *lut,
                    int order, const double *data, int leaf, int s0, int s1, double *lab)
{
    og_grid g = {n_leaves, max_level, B, level, index, lut, order};
    const int L = B + s1 - s0 - 1;
    const int cap = 4096;
    og_term *buf = (og_term *)malloc(sizeof(og_term) * cap);
    const int lam = level[leaf];
    const long B3 = (long)B * B * B;
    int rc = 0;
    for (int z = s0; z < B + s1 - 1; z++)
        for (int y = s0; y < B + s1 - 1; y++)
            for (int x = s0; x < B + s1 - 1; x++) {
                double v;
                if (x >= 0 && x < B && y >= 0 && y < B && z >= 0 && z < B) {
                    v = data[leaf * B3 + x + B * (y + B * z)];
                } else {
                    og_terms T = {buf, 0, cap, 0};
                    const long p[3] = {(long)index[3 * leaf] * B + x, (long)index[3 * leaf + 1] * B + y,
                                       (long)index[3 * leaf + 2] * B + z};
                    og_expand(&g, lam, p, 1.0, &T, 0);
                    if (T.overflow) { rc = -1; v = 0; }
                    else {
                        v = 0.0;
                        for (int k = 0; k < T.n; k++) v += buf[k].w * data[buf[k].leaf * B3 + buf[k].cell];
                    }
                }
                lab[(x - s0) + L * ((y - s0) + L * (z - s0))] = v;
            }
    free(buf);
    return rc;
}

/* FK steps on an adaptive leaf grid, FP64 (Real = double).  fields are
 * [n_leaves][B^3]. h per leaf = 2^-level / B (MRAGrid.inl:191-219).  The lab
 * uses the stencil [-1,2) needed by the RD operator; the operator only reads
 * the 6 face neighbours, so edge/corner ghosts are not evaluated. */
int fk_adaptive_steps_f64(int n_leaves, int max_level, int B, const int *level, const int *index, const int *lut,
                          int order, double *phi, const double *pw, const double *pg,
                          double Dw, double Dg, double rho, double dt, int n_steps, int n_threads)
{
    const long B3 = (long)B * B * B;
    const int L = B + 2;
    const long L3 = (long)L * L * L;
    og_grid g = {n_leaves, max_level, B, level, index, lut, order};
    double *dphidt = (double *)malloc(sizeof(double) * n_leaves * B3);
    int rc = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
    for (int s = 0; s < n_steps; s++) {
#pragma omp parallel
        {
            double *lab = (double *)calloc(3 * L3, sizeof(double));
            const int cap = 4096;
            og_term *buf = (og_term *)malloc(sizeof(og_term) * cap);
#pragma omp for schedule(dynamic, 1)
            for (int b = 0; b < n_leaves; b++) {
                const double *fields[3] = {phi, pw, pg};
                for (int f = 0; f < 3; f++) {
                    double *lf = lab + f * L3;
                    for (int z = -1; z <= B; z++)
                        for (int y = -1; y <= B; y++)
                            for (int x = -1; x <= B; x++) {
                                const int out_x = (x < 0 || x >= B), out_y = (y < 0 || y >= B), out_z = (z < 0 || z >= B);
                                const int nout = out_x + out_y + out_z;
                                double v = 0.0;
                                if (nout == 0) v = fields[f][b * B3 + x + B * (y + B * z)];
                                else if (nout == 1) {
                                    og_terms T = {buf, 0, cap, 0};
                                    const long p[3] = {(long)index[3 * b] * B + x, (long)index[3 * b + 1] * B + y,
                                                       (long)index[3 * b + 2] * B + z};
                                    og_expand(&g, level[b], p, 1.0, &T, 0);
                                    if (T.overflow) rc = -1;
                                    for (int k = 0; k < T.n; k++) v += buf[k].w * fields[f][buf[k].leaf * B3 + buf[k].cell];
                                }
                                lf[(x + 1) + L * ((y + 1) + L * (z + 1))] = v;
                            }
                }
                const double h = ldexp(1.0, -level[b]) / B;
                fk_block_f64(lab, lab + L3, lab + 2 * L3, B, h, Dw, Dg, rho, dphidt + b * B3);
            }
            free(lab);
            free(buf);
        }
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n_leaves * B3; i++) phi[i] = clamp01_f64(phi[i] + dt * dphidt[i]);
    }
    free(dphidt);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* a9: BlockFWT detail, MRAGBlockFWT.h:220-298 + FWTReport :49-192.           */
/* Input: a ghosted lab of phi with stencil [-2,4) for Interp4 (L = B+5) or   */
/* [-1,2) for Interp2 (L = B+2) -- any lab with s0 <= needed low index works; */
/* s0 is the lab's start offset.  Filter per coarse point d and code         */
/* (bx,by,bz): sum over iz,iy,ix of lab(2d-i) * ((wZ*wY) * wX), FP64, loop   */
/* order z, y, x as the reference.  Returns max over codes 1..7 of |value|.   */
/* ------------------------------------------------------------------------ */
static const double fwt_ga2[3] = {-0.5, 1., -0.5};   /* GaSupport {-2,1} */
static const double fwt_ga4[7] = {1 / 16., 0, -9 / 16., 1, -9 / 16., 0, 1 / 16.}; /* {-4,3} */

double fwt_block_detail_f64(const double *lab, int B, int s0, int L, int order)
{
    const double *ga = order == 2 ? fwt_ga2 : fwt_ga4;
    const int gs0 = order == 2 ? -2 : -4, gs1 = order == 2 ? 1 : 3;
    double maxv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int nc = B / 2;
    for (int dz = 0; dz < nc; dz++)
        for (int dy = 0; dy < nc; dy++)
            for (int dx = 0; dx < nc; dx++)
                for (int code = 0; code < 8; code++) {
                    const int bx = code & 1, by = (code >> 1) & 1, bz = (code >> 2) & 1;
                    const int sx = bx ? gs0 : 0, ex = bx ? gs1 : 1;
                    const int sy = by ? gs0 : 0, ey = by ? gs1 : 1;
                    const int sz = bz ? gs0 : 0, ez = bz ? gs1 : 1;
                    double r = 0;
                    for (int iz = sz; iz < ez; iz++) {
                        const double wZ = bz ? ga[iz - gs0] : 1.0;
                        for (int iy = sy; iy < ey; iy++) {
                            const double wY = by ? ga[iy - gs0] : 1.0;
                            const double wZwY = wZ * wY;
                            for (int ix = sx; ix < ex; ix++) {
                                const double wX = bx ? ga[ix - gs0] : 1.0;
                                const int x = 2 * dx - ix - s0, y = 2 * dy - iy - s0, z = 2 * dz - iz - s0;
                                r += lab[x + L * (y + L * z)] * (wZwY * wX);
                            }
                        }
                    }
                    const double a = fabs(r);
                    maxv[code] = maxv[code] > a ? maxv[code] : a; /* std::max(val, maxVal) */
                }
    double m = maxv[1];
    for (int c = 2; c < 8; c++) m = m > maxv[c] ? m : maxv[c];
    return m;
}

/* ------------------------------------------------------------------------ */
/* a16: Glioma_HG_UQ::_dumpUQoutput, Glioma_HG_UQ.cpp:309-372.  Leaves are    */
/* replicated piecewise-constant onto the finest uniform gpd^3 grid; a leaf  */
/* with h > 3*hf writes only a 4^3 patch per cell (the reference's `else`    */
/* branch), so leaves 3+ levels coarser than gpd leave untouched voxels.     */
/* out must be pre-initialised by the caller (zero).                          */
/* ------------------------------------------------------------------------ */
void dump_uniform_f64(int n_leaves, int B, const int *level, const int *index, const double *phi, int gpd, double *out)
{
    const long B3 = (long)B * B * B;
    const double hf = 1. / gpd;
    for (int b = 0; b < n_leaves; b++) {
        const double h = ldexp(1.0, -level[b]) / B;
        const double org[3] = {ldexp((double)index[3 * b], -level[b]), ldexp((double)index[3 * b + 1], -level[b]),
                               ldexp((double)index[3 * b + 2], -level[b])};
        int rep;
        if (h == hf) rep = 1;
        else if (h == 2. * hf) rep = 2;
        else if (h == 3. * hf) rep = 3;
        else rep = 4;
        for (int iz = 0; iz < B; iz++)
            for (int iy = 0; iy < B; iy++)
                for (int ix = 0; ix < B; ix++) {
                    const double x = org[0] + h * ix, y = org[1] + h * iy, z = org[2] + h * iz;
                    const int mx = (int)floor(x / hf), my = (int)floor(y / hf), mz = (int)floor(z / hf);
                    const double v = phi[b * B3 + ix + B * (iy + B * iz)];
                    for (int cz = 0; cz < rep; cz++)
                        for (int cy = 0; cy < rep; cy++)
                            for (int cx = 0; cx < rep; cx++)
                                out[(long)(mx + cx) + gpd * ((long)(my + cy) + (long)gpd * (mz + cz))] = v;
                }
    }
}

/* ------------------------------------------------------------------------ */
/* a17-a19: HGG_Likelihood, Likelihood/source/HGG_Likelihood.cpp:40-189.      */
/* long double accumulation as the reference.                                 */
/* ------------------------------------------------------------------------ */
static int lik_sgn(double d)
{
    const double eps = 0.0;
    if (d < -eps) return -1;
    return d > eps;
}

static long double lik_log_bernoulli(double u, double y, double uc, double slope)
{
    const double is2 = 1. / slope;
    const double diff = u - uc;
    const long double alpha = 0.5 + 0.5 * lik_sgn(diff) * (1. - exp(-diff * diff * is2));
    return (y == 1) ? logl(alpha) : logl(1. - alpha);
}

/* model, pet, t1, t2: nx*ny*nz doubles, x fastest.  points: npts*3 doubles
 * (Sphere.dat D2D layout: column-major, Points(i,c) = pts[i + c*npts]).
 * params: sigma2, petscale, t1uc, t2uc, slope.  Returns log L (not -log L). */
long double lik_ref_f64(const double *model, const double *pet, const double *t1, const double *t2,
                        int nx, int ny, int nz, const double *params, int selective, int step_pet, int step_ti,
                        const double *points, int npts, long double *parts)
{
    const double sigma2 = params[0], petscale = params[1], t1uc = params[2], t2uc = params[3], slope = params[4];
    int N = 0;
    long double sum = 0.;
    for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++)
                if ((ix % step_pet == 0) && (iy % step_pet == 0) && (iz % step_pet == 0)) {
                    const long i = ix + (long)nx * (iy + (long)ny * iz);
                    if (pet[i] > 0.) {
                        sum += (model[i] - petscale * pet[i]) * (model[i] - petscale * pet[i]);
                        N++;
                    }
                }
    const long double p1 = -0.5 * N * log(2. * M_PI * sigma2);
    const long double p2 = -0.5 * (1. / sigma2) * sum;
    const long double lpet = p1 + p2;
    long double lt[2];
    for (int ti = 1; ti <= 2; ti++) {
        const double *data = ti == 1 ? t1 : t2;
        const double uc = ti == 1 ? t1uc : t2uc;
        long double s = 0.;
        if (!selective) {
            for (long i = 0; i < (long)nx * ny * nz; i++) s += lik_log_bernoulli(model[i], data[i], uc, slope);
        } else {
            for (int p = 0; p < npts; p++) {
                const int ix = (int)points[p], iy = (int)points[p + npts], iz = (int)points[p + 2 * npts];
                if ((ix % step_ti == 0) && (iy % step_ti == 0) && (iz % step_ti == 0)) {
                    const long i = ix + (long)nx * (iy + (long)ny * iz);
                    s += lik_log_bernoulli(model[i], data[i], uc, slope);
                }
            }
        }
        lt[ti - 1] = s;
    }
    if (parts) { parts[0] = lpet; parts[1] = lt[0]; parts[2] = lt[1]; }
    return lpet + lt[0] + lt[1];
}

/* Flip-rate Bernoulli + Gaussian PET model (the metric model of the
 * synthetic-observation configs; it has no reference function).  Per voxel:
 *   T1Gd:  m1 = [u > th1]; log P(obs) = log(1-eps) if obs == m1 else log(eps)
 *   FLAIR: m2 = [u > th2]; same
 *   PET:   where an observation exists (pet is not NaN; the generator writes
 *          values only where the ground truth had u > th2):
 *          -0.5*log(2*pi*sigma^2) - (pet - u)^2 / (2*sigma^2)
 * Accumulation in long double, voxel order (x fastest), so that the
 * oracle's own rounding stays far below the device's FP64 tree sums.
 * Returns log L. */
double lik_flip_f64(const double *u, const unsigned char *t1, const unsigned char *t2, const float *pet,
                    long n, double th1, double th2, double eps, double sigma, double *parts)
{
    const double lp1 = log(1. - eps), lp0 = log(eps);
    const double c0 = -0.5 * log(2. * M_PI * sigma * sigma);
    const double is2 = 1. / (2. * sigma * sigma);
    long double s1 = 0, s2 = 0, sp = 0;
    for (long i = 0; i < n; i++) {
        const int m1 = u[i] > th1, m2 = u[i] > th2;
        s1 += (m1 == (int)t1[i]) ? lp1 : lp0;
        s2 += (m2 == (int)t2[i]) ? lp1 : lp0;
        if (!isnan(pet[i])) {
            const double r = (double)pet[i] - u[i];
            sp += c0 - r * r * is2;
        }
    }
    if (parts) { parts[0] = (double)s1; parts[1] = (double)s2; parts[2] = (double)sp; }
    return (double)(s1 + s2 + sp);
}
