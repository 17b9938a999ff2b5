/* Trajectory propagation and correlation functions (test infrastructure).
 *
 * correlation_eval semantics (src/hawaii.c:2116-2347):
 *   c[0] = mu(0).mu(0); for k = 1..L-1: propagate to t_k = k dt, mu(t_k), checks
 *   (NaN dipole -> fail, |c| > 1e100 -> fail, |c_k / c_{k-1}| > 1e10 for k > 1 -> fail),
 *   store c[k], stop after storing once R(t_k) > Rcut.  The backward half restarts from
 *   the initial condition with all momenta inverted (invert_momenta, src/hawaii.c:1171).
 *   crln[k] = 0.5 (c_f[k] + c_b[k]) ALU^3 (src/hawaii.c:2334-2337); a failed check
 *   returns early with crln still zero (src/hawaii.c:2128, 2183-2237), and the caller
 *   accumulates it anyway (src/hawaii.c:3001-3009).
 *
 * Integrator: DOP853 (Hairer, Norsett & Wanner, Solving ODEs I, II.5-II.6), error norm
 * with sk = atol + rtol max(|y0|,|y1|), step controller safe = 0.9, fac1 = 0.333,
 * fac2 = 6, beta = 0, initial step from Hairer's HINIT.  Samples come from the
 * 7th-order dense output, as CVode(CV_NORMAL) interpolates to tout (src/trajectory.c:155).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "../hawaiihybrid_amd/csrc/dop853_coeffs.h"

#define ALU 5.29177210903e-11   /* src/constants.h:8 */
#define NMAX 16

typedef struct {
    const orc_system *sys;
    long nrhs;
} fctx;

static void f(fctx *c, const double *y, double *dy)
{
    c->nrhs++;
    orc_rhs(c->sys, y, dy);
}

static double max2(double a, double b) { return a > b ? a : b; }
static double min2(double a, double b) { return a < b ? a : b; }

/* Hairer HINIT for order 8 */
static double initial_step(fctx *c, int n, const double *y0, const double *f0, double rtol, double atol,
                           double hmax)
{
    double dnf = 0.0, dny = 0.0, der2 = 0.0, h, h1, der12, sk;
    double y1[NMAX], f1[NMAX];
    int i;
    for (i = 0; i < n; ++i) {
        sk = atol + rtol * fabs(y0[i]);
        dnf += (f0[i] / sk) * (f0[i] / sk);
        dny += (y0[i] / sk) * (y0[i] / sk);
    }
    if (dnf <= 1e-10 || dny <= 1e-10) h = 1e-6;
    else h = sqrt(dny / dnf) * 0.01;
    h = min2(h, hmax);
    for (i = 0; i < n; ++i) y1[i] = y0[i] + h * f0[i];
    f(c, y1, f1);
    for (i = 0; i < n; ++i) {
        sk = atol + rtol * fabs(y0[i]);
        der2 += ((f1[i] - f0[i]) / sk) * ((f1[i] - f0[i]) / sk);
    }
    der2 = sqrt(der2) / h;
    der12 = max2(fabs(der2), sqrt(dnf));
    if (der12 <= 1e-15) h1 = max2(1e-6, fabs(h) * 1e-3);
    else h1 = sqrt(sqrt(sqrt(0.01 / der12)));
    return min2(min2(100 * fabs(h), h1), hmax);
}

static void dipole_of(const orc_system *s, const double *y, double mu[3])
{
    double q[8];
    int i, nq = orc_qp_size(s) / 2;
    for (i = 0; i < nq; ++i) q[i] = y[2 * i];
    orc_dipole(s, q, mu);
}

/* vec (optional): the dipole at every recorded sample, vec[3k + i]; cf_checks = 0 keeps only
 * the dipole check (the PR/mu loop, src/hawaii.c:3659-3671) */
static int half_cf(const orc_system *s, const orc_cf_params *p, const double *y0, double *c, double *vec, int *len,
                   orc_traj_stats *st, int cf_checks)
{
    const int n = orc_qp_size(s);
    fctx fc = {s, 0};
    double K[16][NMAX], y[NMAX], ynew[NMAX], ytmp[NMAX], F[7][NMAX], ys[NMAX], ylast[NMAX];
    double mu0[3], mut[3];
    double t = 0.0, h, tend = (p->L - 1) * p->dt, E0;
    double prev = 0.0, curr = 0.0;
    int k = 1, i, j, reject = 0, status = ORC_OK;
    long steps_since_sample = 0;

    memset(c, 0, sizeof(double) * p->L);
    memcpy(y, y0, sizeof(double) * n);
    memcpy(ylast, y0, sizeof(double) * n);
    dipole_of(s, y, mu0);
    st->ndip++;
    c[0] = mu0[0] * mu0[0] + mu0[1] * mu0[1] + mu0[2] * mu0[2];
    if (vec) {
        memset(vec, 0, sizeof(double) * 3 * p->L);
        vec[0] = mu0[0];
        vec[1] = mu0[1];
        vec[2] = mu0[2];
    }
    *len = 1;
    E0 = orc_hamiltonian(s, y0);
    if (p->L <= 1) return ORC_OK;

    f(&fc, y, K[0]);
    h = initial_step(&fc, n, y, K[0], p->rtol, p->atol, tend);

    for (;;) {
        double err = 0.0, err2 = 0.0, fac11, fac, hnew;
        /* CVode's mxstep test: max_steps accepted steps without reaching the next output
         * time ends the call with CV_TOO_MUCH_WORK (src/trajectory.c:15, 62) */
        if (steps_since_sample >= p->max_steps) { status = ORC_FAIL_STEPS; break; }
        if (0.1 * fabs(h) <= fabs(t) * 2.220446049250313e-16) { status = ORC_FAIL_STEPSIZE; break; }
        if ((t + 1.01 * h - tend) > 0.0) h = tend - t;

        /* stages 1..11 */
        for (j = 1; j < 12; ++j) {
            for (i = 0; i < n; ++i) {
                double acc = 0.0;
                int m;
                for (m = 0; m < j; ++m) acc += DOP853_A[j][m] * K[m][i];
                ytmp[i] = y[i] + h * acc;
            }
            f(&fc, ytmp, K[j]);
        }
        for (i = 0; i < n; ++i) {
            double acc = 0.0;
            for (j = 0; j < 12; ++j) acc += DOP853_B[j] * K[j][i];
            ynew[i] = y[i] + h * acc;
        }
        f(&fc, ynew, K[12]);

        for (i = 0; i < n; ++i) {
            double sk = p->atol + p->rtol * max2(fabs(y[i]), fabs(ynew[i]));
            double e5 = 0.0, e3 = 0.0;
            for (j = 0; j < 13; ++j) {
                e5 += DOP853_E5[j] * K[j][i];
                e3 += DOP853_E3[j] * K[j][i];
            }
            err += (e5 / sk) * (e5 / sk);
            err2 += (e3 / sk) * (e3 / sk);
        }
        {
            double deno = err + 0.01 * err2;
            if (deno <= 0.0) deno = 1.0;
            err = fabs(h) * err * sqrt(1.0 / (n * deno));
        }
        fac11 = sqrt(sqrt(sqrt(err)));   /* err^(1/8) */
        fac = max2(1.0 / 6.0, min2(1.0 / 0.333, fac11 / 0.9));
        hnew = h / fac;

        if (err > 1.0) {
            hnew = h / min2(1.0 / 0.333, fac11 / 0.9);
            reject = 1;
            st->nreject++;
            h = hnew;
            continue;
        }

        /* accepted step [t, t+h] */
        st->naccept++;
        steps_since_sample++;
        {
            double tnew = t + h;
            int have_dense = 0;
            while (k < p->L && k * p->dt <= tnew) {
                double tk = k * p->dt;
                if (tk == tnew) {
                    memcpy(ys, ynew, sizeof(double) * n);
                } else {
                    double x, omx;
                    if (!have_dense) {
                        for (j = 13; j < 16; ++j) {
                            for (i = 0; i < n; ++i) {
                                double acc = 0.0;
                                int m;
                                for (m = 0; m < j; ++m) acc += DOP853_A[j][m] * K[m][i];
                                ytmp[i] = y[i] + h * acc;
                            }
                            f(&fc, ytmp, K[j]);
                        }
                        for (i = 0; i < n; ++i) {
                            double dy = ynew[i] - y[i];
                            F[0][i] = dy;
                            F[1][i] = h * K[0][i] - dy;
                            F[2][i] = 2.0 * dy - h * (K[12][i] + K[0][i]);
                            for (j = 0; j < 4; ++j) {
                                double acc = 0.0;
                                int m;
                                for (m = 0; m < 16; ++m) acc += DOP853_D[j][m] * K[m][i];
                                F[3 + j][i] = h * acc;
                            }
                        }
                        have_dense = 1;
                    }
                    x = (tk - t) / h;
                    omx = 1.0 - x;
                    for (i = 0; i < n; ++i) {
                        /* y_old + x(F0 + (1-x)(F1 + x(F2 + (1-x)(F3 + x(F4 + (1-x)(F5 + x F6)))))) */
                        double v = F[5][i] + x * F[6][i];
                        v = F[4][i] + omx * v;
                        v = F[3][i] + x * v;
                        v = F[2][i] + omx * v;
                        v = F[1][i] + x * v;
                        v = F[0][i] + omx * v;
                        ys[i] = y[i] + x * v;
                    }
                }
                dipole_of(s, ys, mut);
                st->ndip++;
                if (isnan(mut[0]) || isnan(mut[1]) || isnan(mut[2])) {
                    if (cf_checks) { status = ORC_FAIL_DIPOLE_NAN; goto done; }
                    /* PR/mu: the reference skips the sample and integrates on (src/hawaii.c:
                     * 3664-3677); it stays 0 in its zeroed series */
                    if (vec) vec[3 * k] = vec[3 * k + 1] = vec[3 * k + 2] = 0.0;
                    *len = k + 1;
                    memcpy(ylast, ys, sizeof(double) * n);
                    steps_since_sample = 0;
                    ++k;
                    continue;
                }
                prev = curr;
                curr = mu0[0] * mut[0] + mu0[1] * mut[1] + mu0[2] * mut[2];
                if (cf_checks && fabs(curr) > 1e100) { status = ORC_FAIL_CF_HUGE; goto done; }
                if (cf_checks && k > 1 && fabs(curr / prev) > 1e10) { status = ORC_FAIL_CF_JUMP; goto done; }
                c[k] = curr;
                if (vec) {
                    vec[3 * k] = mut[0];
                    vec[3 * k + 1] = mut[1];
                    vec[3 * k + 2] = mut[2];
                }
                *len = k + 1;
                memcpy(ylast, ys, sizeof(double) * n);
                steps_since_sample = 0;
                if (ys[4] > p->Rcut) goto done;
                ++k;
            }
            if (k >= p->L) goto done;
            memcpy(y, ynew, sizeof(double) * n);
            memcpy(K[0], K[12], sizeof(double) * n);
            t = tnew;
        }
        if (fabs(hnew) > tend) hnew = tend;   /* hmax = whole interval (DOP853 default) */
        if (reject) { hnew = min2(fabs(hnew), fabs(h)); reject = 0; }
        h = hnew;
    }
done:
    st->nrhs += fc.nrhs;
    {
        /* energy at the last recorded sample (make_step's check, src/trajectory.c:167-180) */
        double Eend = orc_hamiltonian(s, ylast);
        if (fabs(Eend - E0) > st->e_drift) st->e_drift = fabs(Eend - E0);
    }
    return status;
}

int orc_half_cf(const orc_system *s, const orc_cf_params *p, const double *y0, double *c, int *len,
                orc_traj_stats *st)
{
    return half_cf(s, p, y0, c, NULL, len, st, 1);
}

/* The PR/mu trajectory (src/hawaii.c:3578-3672): from y0 forward, the dipole recorded at
 * t = 0, dt, 2 dt, ... up to and including the first sample with R > p->Rcut, at most p->L
 * samples; vec[3k + i] holds component i at t = k dt (the reference's series is k >= 1). */
int orc_prmu_series(const orc_system *s, const orc_cf_params *p, const double *y0, double *vec, int *len,
                    orc_traj_stats *st)
{
    double *c = calloc(p->L, sizeof(double));
    int status = half_cf(s, p, y0, c, vec, len, st, 0);
    free(c);
    return status;
}

void orc_zimmermann_series_cf(const double *d, int L, double *crln)
{
    int sh, c, i;
    for (sh = 0; sh < L; ++sh) {
        double acc = 0.0;
        for (c = 0; c < L; ++c)
            for (i = 0; i < 3; ++i) acc += d[3 * c + i] * d[3 * (c + sh) + i] * ALU * ALU * ALU;
        crln[sh] = acc / L;
    }
}

/* correlation_eval_zimmerman_trick (src/hawaii.c:1888-2114): the series d[0 .. 2L-2] holds the
 * backward half reversed below mu(0) at d[L-1] and the forward half above it; a dipole failure
 * in either half returns with crln zeroed (the forward one before the backward half runs); a
 * step failure ends that half and the status of the backward half is the one returned. */
static int traj_cf_zimmermann(const orc_system *s, const orc_cf_params *p, const double *y0, double *crln,
                              int *len_f, int *len_b, orc_traj_stats *st)
{
    double *c = calloc(p->L, sizeof(double)), *vf = calloc(3 * (size_t)p->L, sizeof(double));
    double *vb = calloc(3 * (size_t)p->L, sizeof(double)), *d = calloc(3 * (size_t)(2 * p->L - 1), sizeof(double));
    double yb[NMAX];
    int n = orc_qp_size(s), i, k, status;
    *len_b = 0;
    memset(crln, 0, sizeof(double) * p->L);
    status = half_cf(s, p, y0, c, vf, len_f, st, 1);
    if (status == ORC_FAIL_DIPOLE_NAN || status == ORC_FAIL_CF_HUGE || status == ORC_FAIL_CF_JUMP) goto out;
    for (i = 0; i < n; ++i) yb[i] = (i % 2 == 1) ? -y0[i] : y0[i];
    status = half_cf(s, p, yb, c, vb, len_b, st, 1);
    if (status == ORC_FAIL_DIPOLE_NAN || status == ORC_FAIL_CF_HUGE || status == ORC_FAIL_CF_JUMP) goto out;
    if (status == ORC_FAIL_STEPS) goto out;
    for (k = 0; k < *len_f; ++k)
        for (i = 0; i < 3; ++i) d[3 * (p->L - 1 + k) + i] = vf[3 * k + i];
    for (k = 1; k < *len_b; ++k)
        for (i = 0; i < 3; ++i) d[3 * (p->L - 1 - k) + i] = vb[3 * k + i];
    orc_zimmermann_series_cf(d, p->L, crln);
out:
    free(c);
    free(vf);
    free(vb);
    free(d);
    return status;
}

int orc_traj_cf(const orc_system *s, const orc_cf_params *p, const double *y0, double *crln,
                int *len_f, int *len_b, orc_traj_stats *st)
{
    if (p->zimmermann) return traj_cf_zimmermann(s, p, y0, crln, len_f, len_b, st);
    {
    double *cf = calloc(p->L, sizeof(double)), *cb = calloc(p->L, sizeof(double));
    double yb[NMAX];
    int n = orc_qp_size(s), i, status;
    *len_b = 0;
    memset(crln, 0, sizeof(double) * p->L);
    status = orc_half_cf(s, p, y0, cf, len_f, st);
    if (status == ORC_FAIL_DIPOLE_NAN || status == ORC_FAIL_CF_HUGE || status == ORC_FAIL_CF_JUMP) goto out;
    if (p->halves == ORC_HALVES_BOTH) {
        for (i = 0; i < n; ++i) yb[i] = (i % 2 == 1) ? -y0[i] : y0[i];
        status = orc_half_cf(s, p, yb, cb, len_b, st);
        if (status == ORC_FAIL_DIPOLE_NAN || status == ORC_FAIL_CF_HUGE || status == ORC_FAIL_CF_JUMP) goto out;
    }
    /* The caller drops a trajectory whose last integrator call ended on the step limit
     * (CVode's CV_TOO_MUCH_WORK = -1 returned by correlation_eval, `if (status == -1)
     * continue;` src/hawaii.c:3001); it still counts in ntraj (src/hawaii.c:3040). */
    if (status == ORC_FAIL_STEPS) goto out;
    if (p->halves == ORC_HALVES_BOTH) {
        for (i = 0; i < p->L; ++i) crln[i] = 0.5 * (cf[i] + cb[i]) * ALU * ALU * ALU;
    } else {
        for (i = 0; i < p->L; ++i) crln[i] = cf[i] * ALU * ALU * ALU;
    }
out:
    free(cf);
    free(cb);
    return status;
    }
}

typedef struct {
    const orc_system *s;
    const orc_cf_params *p;
    const double *ic;
    long ntraj;
    int nthreads, tid;
    double *crln;
    int *status, *len_f, *len_b;
    long *nrhs, *ndip;
    double *e_drift;
    double *partial;   /* ensemble mode */
    double ppf;
    long nrhs_sum;
} job;

static void *rows_worker(void *arg)
{
    job *jb = arg;
    int n = orc_qp_size(jb->s);
    long t;
    double *tmp = jb->partial ? calloc(jb->p->L, sizeof(double)) : NULL;
    for (t = jb->tid; t < jb->ntraj; t += jb->nthreads) {
        orc_traj_stats st = {0, 0, 0, 0, 0.0};
        int lf, lb, status;
        double *row = jb->partial ? tmp : jb->crln + (size_t)t * jb->p->L;
        status = orc_traj_cf(jb->s, jb->p, jb->ic + (size_t)t * n, row, &lf, &lb, &st);
        if (jb->partial) {
            int k;
            for (k = 0; k < jb->p->L; ++k) jb->partial[k] += jb->ppf * row[k];
            jb->nrhs_sum += st.nrhs;
        } else {
            jb->status[t] = status;
            jb->len_f[t] = lf;
            jb->len_b[t] = lb;
            jb->nrhs[t] = st.nrhs;
            jb->ndip[t] = st.ndip;
            jb->e_drift[t] = st.e_drift;
        }
    }
    free(tmp);
    return NULL;
}

void orc_cf_rows(const orc_system *s, const orc_cf_params *p, const double *ic, long ntraj, int nthreads,
                 double *crln, int *status, int *len_f, int *len_b, long *nrhs, long *ndip, double *e_drift)
{
    pthread_t th[256];
    job jobs[256];
    int i;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    for (i = 0; i < nthreads; ++i) {
        job jb = {s, p, ic, ntraj, nthreads, i, crln, status, len_f, len_b, nrhs, ndip, e_drift, NULL, 1.0, 0};
        jobs[i] = jb;
        pthread_create(&th[i], NULL, rows_worker, &jobs[i]);
    }
    for (i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
}

void orc_cf_ensemble(const orc_system *s, const orc_cf_params *p, const double *ic, long ntraj, int nthreads,
                     double ppf, double *cf, long *nrhs_total)
{
    pthread_t th[256];
    job jobs[256];
    int i, k;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    for (i = 0; i < nthreads; ++i) {
        job jb = {s, p, ic, ntraj, nthreads, i, NULL, NULL, NULL, NULL, NULL, NULL, NULL,
                  calloc(p->L, sizeof(double)), ppf, 0};
        jobs[i] = jb;
        pthread_create(&th[i], NULL, rows_worker, &jobs[i]);
    }
    for (k = 0; k < p->L; ++k) cf[k] = 0.0;
    if (nrhs_total) *nrhs_total = 0;
    for (i = 0; i < nthreads; ++i) {
        pthread_join(th[i], NULL);
        for (k = 0; k < p->L; ++k) cf[k] += jobs[i].partial[k];
        if (nrhs_total) *nrhs_total += jobs[i].nrhs_sum;
        free(jobs[i].partial);
    }
    for (k = 0; k < p->L; ++k) cf[k] /= (double)ntraj;
}
