/*
 * adjoint_oracle.c -- CPU oracle of the discrete adjoint (TEST INFRASTRUCTURE
 * ONLY; only tests/ use it).
 *
 * Forward: the reference's explicit FK step (Case2 operator + UpdateTumor,
 * Glioma/Glioma_ReactionDiffusionOperator.h:148-266) in FP64 on a uniform
 * periodic N^3 grid, dt and the step count fixed.  The reference has no
 * adjoint; this is the exact reverse-mode derivative of that discrete map:
 *
 *   v_c  = u_c + dt*(ih2*sum_k df_ck u_k - ih2*(sum_k df_ck) u_c + rho u_c(1-u_c))
 *   u'_c = clamp01(v_c)            d u'_c / d v_c = [0 < v_c < 1]  (NaN -> 0)
 *
 * With mu = mask .* lambda_{n+1}:
 *   lambda_n(c) = mu_c (1 + dt(rho(1-2u_c) - ih2 sum_k df_ck)) + dt ih2 sum_k df_kc mu_k
 *   dL/dlogDw  += sum_c mu_c dt ih2 (sum_k df_ck u_k - (sum_k df_ck) u_c)   (df ~ Dw, Dg = Dw/10)
 *   dL/dlogrho += sum_c mu_c dt rho u_c (1-u_c)
 * Terms whose coefficient is not finite multiply a zero adjoint (the cells
 * next to CSF are clamped to 0 by the NaN quirk) and are skipped.
 *
 * Pinned by central finite differences of the forward map in tests.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline long wrapl(long i, long n)
{
    i %= n;
    return i < 0 ? i + n : i;
}

/* per-cell face coefficients of the Case2 operator (with the Neumann doubling) */
static void face_coeffs(const double *phi, const double *pw, const double *pg, long N, double Dw, double Dg,
                        double *df /* N^3 x 6 */)
{
    const long N3 = N * N * N;
    for (long c = 0; c < N3; c++) {
        const long x = c % N, y = (c / N) % N, z = c / (N * N);
        long nb[6];
        nb[0] = wrapl(x - 1, N) + N * (y + N * z);
        nb[1] = wrapl(x + 1, N) + N * (y + N * z);
        nb[2] = x + N * (wrapl(y - 1, N) + N * z);
        nb[3] = x + N * (wrapl(y + 1, N) + N * z);
        nb[4] = x + N * (y + N * wrapl(z - 1, N));
        nb[5] = x + N * (y + N * wrapl(z + 1, N));
        double d[6] = {0, 0, 0, 0, 0, 0};
        const double tissue = pw[c] + pg[c] + phi[c];
        if (tissue > 0.) {
            const double b0 = pw[c] + pg[c];
            for (int k = 0; k < 6; k++) {
                const double bk = pw[nb[k]] + pg[nb[k]];
                d[k] = 2.0 * (1.0 / ((1.0 / ((pw[c] * Dw + pg[c] * Dg) / b0)) + (1.0 / ((pw[nb[k]] * Dw + pg[nb[k]] * Dg) / bk))));
            }
        }
        for (int k = 0; k < 6; k += 2) {
            if (d[k] == 0 && phi[nb[k]] > 0.) d[k + 1] *= 2.0;
            if (d[k + 1] == 0 && phi[nb[k + 1]] > 0.) d[k] *= 2.0;
        }
        memcpy(df + 6 * c, d, sizeof(d));
    }
}

static void neighbours(long c, long N, long nb[6])
{
    const long x = c % N, y = (c / N) % N, z = c / (N * N);
    nb[0] = wrapl(x - 1, N) + N * (y + N * z);
    nb[1] = wrapl(x + 1, N) + N * (y + N * z);
    nb[2] = x + N * (wrapl(y - 1, N) + N * z);
    nb[3] = x + N * (wrapl(y + 1, N) + N * z);
    nb[4] = x + N * (y + N * wrapl(z - 1, N));
    nb[5] = x + N * (y + N * wrapl(z + 1, N));
}

/* Forward n_steps from u0 (stores every state), then the adjoint of the
 * terminal cotangent dLduN.  Outputs: uN, dLdu0, grads[0] = dL/dlogDw,
 * grads[1] = dL/dlogrho.  Returns 0, -1 on allocation failure. */
int adj_fk_f64(const double *pw, const double *pg, long N, double h, double Dw, double Dg, double rho, double dt,
               int n_steps, const double *u0, const double *dLduN, double *uN, double *dLdu0, double *grads)
{
    const long N3 = N * N * N;
    const double ih2 = 1. / (h * h);
    double *U = (double *)malloc(sizeof(double) * N3 * (size_t)(n_steps + 1));
    unsigned char *M = (unsigned char *)malloc((size_t)N3 * n_steps);
    double *df = (double *)malloc(sizeof(double) * 6 * N3);
    double *lam = (double *)malloc(sizeof(double) * N3);
    double *mu = (double *)malloc(sizeof(double) * N3);
    if (!U || !M || !df || !lam || !mu) {
        free(U); free(M); free(df); free(lam); free(mu);
        return -1;
    }
    memcpy(U, u0, sizeof(double) * N3);
    for (int n = 0; n < n_steps; n++) {
        const double *u = U + (size_t)n * N3;
        double *un = U + (size_t)(n + 1) * N3;
        face_coeffs(u, pw, pg, N, Dw, Dg, df);
        for (long c = 0; c < N3; c++) {
            long nb[6];
            neighbours(c, N, nb);
            const double *d = df + 6 * c;
            const double in = ih2 * (d[0] * u[nb[0]] + d[1] * u[nb[1]] + d[2] * u[nb[2]] + d[3] * u[nb[3]] +
                                     d[4] * u[nb[4]] + d[5] * u[nb[5]]);
            const double out = -((d[0] + d[1] + d[2] + d[3] + d[4] + d[5]) * u[c] * ih2);
            const double reac = rho * u[c] * (1. - u[c]);
            const double v = u[c] + dt * (out + in + reac);
            M[(size_t)n * N3 + c] = (v > 0. && v < 1.) ? 1 : 0;
            double w = (0. < v) ? v : 0.;
            un[c] = (w < 1.) ? w : 1.;
        }
    }
    memcpy(uN, U + (size_t)n_steps * N3, sizeof(double) * N3);
    memcpy(lam, dLduN, sizeof(double) * N3);
    double gD = 0, gR = 0;
    for (int n = n_steps - 1; n >= 0; n--) {
        const double *u = U + (size_t)n * N3;
        face_coeffs(u, pw, pg, N, Dw, Dg, df);
        for (long c = 0; c < N3; c++) mu[c] = M[(size_t)n * N3 + c] ? lam[c] : 0.0;
        for (long c = 0; c < N3; c++) {
            long nb[6];
            neighbours(c, N, nb);
            const double *d = df + 6 * c;
            double acc = 0.0;
            if (mu[c] != 0.0) {
                const double sd = d[0] + d[1] + d[2] + d[3] + d[4] + d[5];
                acc = mu[c] * (1. + dt * (rho * (1. - 2. * u[c]) - ih2 * sd));
                const double diff = ih2 * (d[0] * u[nb[0]] + d[1] * u[nb[1]] + d[2] * u[nb[2]] + d[3] * u[nb[3]] +
                                           d[4] * u[nb[4]] + d[5] * u[nb[5]]) - sd * u[c] * ih2;
                gD += mu[c] * dt * diff;
                gR += mu[c] * dt * rho * u[c] * (1. - u[c]);
            }
            /* transpose: contributions of the neighbours' updates; neighbour
             * nb[k] reaches c through its face opposite to k */
            static const int opp[6] = {1, 0, 3, 2, 5, 4};
            for (int k = 0; k < 6; k++) {
                const long q = nb[k];
                if (mu[q] == 0.0) continue;
                const double dq = df[6 * q + opp[k]];
                if (isfinite(dq)) acc += dt * ih2 * dq * mu[q];
            }
            lam[c] = acc;
        }
    }
    memcpy(dLdu0, lam, sizeof(double) * N3);
    grads[0] = gD;
    grads[1] = gR;
    free(U); free(M); free(df); free(lam); free(mu);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* The same forward map and adjoint with checkpoints (every ckpt_every steps) */
/* and recomputation of each segment in the backward pass, OpenMP over cells, */
/* for grids and step counts whose stored states would not fit in memory     */
/* (128^3 x 2000 steps: 33 GB).  Per cell the arithmetic is that of          */
/* adj_fk_f64; the gradient sums are OpenMP reductions (their order differs). */
/* ------------------------------------------------------------------------ */
#ifdef _OPENMP
#include <omp.h>
#endif

static void fwd_step(const double *u, double *un, unsigned char *M, const double *df, long N, double ih2, double rho,
                     double dt)
{
    const long N3 = N * N * N;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < N3; c++) {
        long nb[6];
        neighbours(c, N, nb);
        const double *d = df + 6 * c;
        const double in = ih2 * (d[0] * u[nb[0]] + d[1] * u[nb[1]] + d[2] * u[nb[2]] + d[3] * u[nb[3]] +
                                 d[4] * u[nb[4]] + d[5] * u[nb[5]]);
        const double out = -((d[0] + d[1] + d[2] + d[3] + d[4] + d[5]) * u[c] * ih2);
        const double reac = rho * u[c] * (1. - u[c]);
        const double v = u[c] + dt * (out + in + reac);
        if (M) M[c] = (v > 0. && v < 1.) ? 1 : 0;
        double w = (0. < v) ? v : 0.;
        un[c] = (w < 1.) ? w : 1.;
    }
}

/* The harmonic-mean coefficients of face_coeffs depend on phi only through
 * the tissue test (pw+pg+phi > 0) and the Neumann doubling: base[] holds them
 * for cells with pw+pg > 0 (where the test always passes); a cell with
 * pw+pg == 0 takes 0/0 = NaN coefficients exactly when phi > 0 (as
 * face_coeffs computes them), and 0 otherwise. */
static void face_base(const double *pw, const double *pg, long N, double Dw, double Dg, double *base)
{
    const long N3 = N * N * N;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < N3; c++) {
        long nb[6];
        neighbours(c, N, nb);
        const double b0 = pw[c] + pg[c];
        for (int k = 0; k < 6; k++) {
            const double bk = pw[nb[k]] + pg[nb[k]];
            base[6 * c + k] = 2.0 * (1.0 / ((1.0 / ((pw[c] * Dw + pg[c] * Dg) / b0)) +
                                            (1.0 / ((pw[nb[k]] * Dw + pg[nb[k]] * Dg) / bk))));
        }
    }
}

static void face_coeffs_par(const double *phi, const double *pw, const double *pg, const double *base, long N,
                            double *df)
{
    const long N3 = N * N * N;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < N3; c++) {
        long nb[6];
        neighbours(c, N, nb);
        double d[6] = {0, 0, 0, 0, 0, 0};
        const double tissue = pw[c] + pg[c] + phi[c];
        if (tissue > 0.)
            for (int k = 0; k < 6; k++) d[k] = base[6 * c + k];
        for (int k = 0; k < 6; k += 2) {
            if (d[k] == 0 && phi[nb[k]] > 0.) d[k + 1] *= 2.0;
            if (d[k + 1] == 0 && phi[nb[k + 1]] > 0.) d[k] *= 2.0;
        }
        memcpy(df + 6 * c, d, sizeof(d));
    }
}

/* Forward n_steps from u0 (keeping every ckpt_every-th state), uN; then the
 * adjoint of the terminal cotangent -- dLduN if not null, else that of the
 * Gaussian PET term (pet - uN)/sigma^2 on the voxels where pet is not NaN,
 * when pet is not null: dLdu0 and grads as adj_fk_f64.  Returns 0, -1 on
 * allocation failure. */
int adj_fk_f64_ckpt(const double *pw, const double *pg, long N, double h, double Dw, double Dg, double rho, double dt,
                    int n_steps, int ckpt_every, int n_threads, const double *u0, const double *dLduN,
                    const float *pet, double sigma, double *uN, double *dLdu0, double *grads)
{
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
    const long N3 = N * N * N;
    const double ih2 = 1. / (h * h);
    const int K = ckpt_every > 0 ? ckpt_every : 1;
    const int nck = (n_steps + K - 1) / K + 1;
    double *CK = (double *)malloc(sizeof(double) * N3 * (size_t)nck);
    double *S = (double *)malloc(sizeof(double) * N3 * (size_t)(K + 1));
    unsigned char *M = (unsigned char *)malloc((size_t)N3 * K);
    double *df = (double *)malloc(sizeof(double) * 6 * N3);
    double *lam = (double *)malloc(sizeof(double) * N3);
    double *mu = (double *)malloc(sizeof(double) * N3);
    double *tmp = (double *)malloc(sizeof(double) * N3);
    double *base = (double *)malloc(sizeof(double) * 6 * N3);
    if (!CK || !S || !M || !df || !lam || !mu || !tmp || !base) {
        free(CK); free(S); free(M); free(df); free(lam); free(mu); free(tmp); free(base);
        return -1;
    }
    face_base(pw, pg, N, Dw, Dg, base);
    memcpy(CK, u0, sizeof(double) * N3);
    memcpy(tmp, u0, sizeof(double) * N3);
    for (int n = 0; n < n_steps; n++) {
        face_coeffs_par(tmp, pw, pg, base, N, df);
        fwd_step(tmp, lam, NULL, df, N, ih2, rho, dt);
        memcpy(tmp, lam, sizeof(double) * N3);
        if ((n + 1) % K == 0) memcpy(CK + (size_t)((n + 1) / K) * N3, tmp, sizeof(double) * N3);
    }
    memcpy(uN, tmp, sizeof(double) * N3);
    if (!dLduN && !pet) {
        free(CK); free(S); free(M); free(df); free(lam); free(mu); free(tmp); free(base);
        return 0;
    }
    if (dLduN) memcpy(lam, dLduN, sizeof(double) * N3);
    else
        for (long c = 0; c < N3; c++) lam[c] = isnan(pet[c]) ? 0.0 : ((double)pet[c] - uN[c]) / (sigma * sigma);
    double gD = 0, gR = 0;
    for (int seg = (n_steps - 1) / K; seg >= 0; seg--) {
        const int n0 = seg * K, n1 = (n0 + K < n_steps) ? n0 + K : n_steps;
        memcpy(S, CK + (size_t)seg * N3, sizeof(double) * N3);
        for (int n = n0; n < n1; n++) {
            double *u = S + (size_t)(n - n0) * N3;
            face_coeffs_par(u, pw, pg, base, N, df);
            fwd_step(u, u + N3, M + (size_t)(n - n0) * N3, df, N, ih2, rho, dt);
        }
        for (int n = n1 - 1; n >= n0; n--) {
            const double *u = S + (size_t)(n - n0) * N3;
            const unsigned char *Mn = M + (size_t)(n - n0) * N3;
            face_coeffs_par(u, pw, pg, base, N, df);
#pragma omp parallel for schedule(static)
            for (long c = 0; c < N3; c++) mu[c] = Mn[c] ? lam[c] : 0.0;
#pragma omp parallel for schedule(static) reduction(+ : gD, gR)
            for (long c = 0; c < N3; c++) {
                long nb[6];
                neighbours(c, N, nb);
                const double *d = df + 6 * c;
                double acc = 0.0;
                if (mu[c] != 0.0) {
                    const double sd = d[0] + d[1] + d[2] + d[3] + d[4] + d[5];
                    acc = mu[c] * (1. + dt * (rho * (1. - 2. * u[c]) - ih2 * sd));
                    const double diff = ih2 * (d[0] * u[nb[0]] + d[1] * u[nb[1]] + d[2] * u[nb[2]] +
                                               d[3] * u[nb[3]] + d[4] * u[nb[4]] + d[5] * u[nb[5]]) -
                                        sd * u[c] * ih2;
                    gD += mu[c] * dt * diff;
                    gR += mu[c] * dt * rho * u[c] * (1. - u[c]);
                }
                static const int opp[6] = {1, 0, 3, 2, 5, 4};
                for (int k = 0; k < 6; k++) {
                    const long q = nb[k];
                    if (mu[q] == 0.0) continue;
                    const double dq = df[6 * q + opp[k]];
                    if (isfinite(dq)) acc += dt * ih2 * dq * mu[q];
                }
                tmp[c] = acc;
            }
            memcpy(lam, tmp, sizeof(double) * N3);
        }
    }
    memcpy(dLdu0, lam, sizeof(double) * N3);
    grads[0] = gD;
    grads[1] = gR;
    free(CK); free(S); free(M); free(df); free(lam); free(mu); free(tmp); free(base);
    return 0;
}
