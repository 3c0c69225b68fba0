/*
 * CPU float64 oracle for the 2D ring-array FDTD hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: only tests/, the smoke
 * entry point and the benchmark's CPU-baseline leg load it. The product path
 * (soundwavesimulation_amd) never links, loads or calls it.
 *
 * It restates, in plain C with double arithmetic, the discrete scheme of the
 * reference's batched torch path:
 *   - leapfrog interior  u' = 2u - u_old + dt^2 c^2 Lap(u)   (zero-padded 5-point
 *     Laplacian, as conv2d(padding=1)):  fdtd2d/torch_backend.py:493-522,
 *     fdtd2d/solver.py:52-58, fdtd2d/wave.py:109-129;
 *   - soft point source added before the boundary:  torch_backend.py:517;
 *   - first-order Mur edges, then corners averaged from an x- and a y-update that
 *     read the freshly updated edge cells:  torch_backend.py:477-490,
 *     solver.py:90-124;  coefficients k = (c dt - h)/(c dt + h):
 *     torch_backend.py:264-275;
 *   - bilinear receiver gather after the boundary:  torch_backend.py:540-560;
 *   - adjoint: residual = pred - obs with the transmitter channel zeroed
 *     (invert_ring.py:603-609), bilinear scatter (torch_backend.py:563-580),
 *     L*[lam] = Lap(c^2 lam) with the *forward* Mur operator
 *     (torch_backend.py:525-537), imaging g += K dt^2 lam' * Lap(u_fwd[n-1]) on
 *     steps with (n-1) % K == 0 (invert_ring.py:683-701, torch_backend.py:583-598);
 *   - misfit J = 1/2 sum residual^2 (invert_ring.py:651);
 *   - optional bilinear transmitters (the reference's suggested reciprocity
 *     fix, desmond/docs/upstream_suggestions.md:65-71, applied in
 *     desmond/src/phase2_numerics.py:105-120): w * value added on the four
 *     footprint nodes with the receivers' bilinear weights, in place of the
 *     nearest-node point source.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(y, x) ((size_t)(y) * (size_t)nx + (size_t)(x))

typedef struct {
    int ny, nx;
    double dt2, idx2, idy2, cen;
    const double *c2;
    double *k_left, *k_right, *k_bottom, *k_top; /* indexed by row / column */
    double kx[4], ky[4];                          /* corners (0,0),(0,-1),(-1,0),(-1,-1) */
} grid_t;

static void grid_init(grid_t *g, int ny, int nx, double dx, double dy, double dt,
                      const double *c, double *c2_store, double *k_store)
{
    g->ny = ny;
    g->nx = nx;
    g->dt2 = dt * dt;
    g->idx2 = 1.0 / (dx * dx);
    g->idy2 = 1.0 / (dy * dy);
    g->cen = -2.0 * (g->idx2 + g->idy2);
    for (size_t i = 0; i < (size_t)ny * nx; ++i) c2_store[i] = c[i] * c[i];
    g->c2 = c2_store;
    g->k_left = k_store;
    g->k_right = k_store + ny;
    g->k_bottom = k_store + 2 * ny;
    g->k_top = k_store + 2 * ny + nx;
    for (int y = 0; y < ny; ++y) {
        double cl = c[IDX(y, 0)], cr = c[IDX(y, nx - 1)];
        g->k_left[y] = (cl * dt - dx) / (cl * dt + dx);
        g->k_right[y] = (cr * dt - dx) / (cr * dt + dx);
    }
    for (int x = 0; x < nx; ++x) {
        double cb = c[IDX(0, x)], ct = c[IDX(ny - 1, x)];
        g->k_bottom[x] = (cb * dt - dy) / (cb * dt + dy);
        g->k_top[x] = (ct * dt - dy) / (ct * dt + dy);
    }
    const int cr[4] = {0, 0, ny - 1, ny - 1}, cc[4] = {0, nx - 1, 0, nx - 1};
    for (int i = 0; i < 4; ++i) {
        double cv = c[IDX(cr[i], cc[i])];
        g->kx[i] = (cv * dt - dx) / (cv * dt + dx);
        g->ky[i] = (cv * dt - dy) / (cv * dt + dy);
    }
}

/* Zero-padded 5-point Laplacian of f at (y, x). */
static inline double lap_at(const grid_t *g, const double *f, int y, int x)
{
    const int nx = g->nx, ny = g->ny;
    double w = x > 0 ? f[IDX(y, x - 1)] : 0.0;
    double e = x < nx - 1 ? f[IDX(y, x + 1)] : 0.0;
    double s = y > 0 ? f[IDX(y - 1, x)] : 0.0;
    double n = y < ny - 1 ? f[IDX(y + 1, x)] : 0.0;
    return g->idx2 * (w + e) + g->idy2 * (s + n) + g->cen * f[IDX(y, x)];
}

/* Interior leapfrog into un over the whole grid (boundary cells are fixed up
 * by mur()). adjoint != 0 uses Lap(c^2 f) with work = c^2 f precomputed. */
static void leapfrog(const grid_t *g, const double *uo, const double *u, double *un,
                     int adjoint, double *work)
{
    const int ny = g->ny, nx = g->nx;
    const size_t n = (size_t)ny * nx;
    if (adjoint) {
#pragma omp parallel for schedule(static) if (n > 65536)
        for (long i = 0; i < (long)n; ++i) work[i] = g->c2[i] * u[i];
    }
#pragma omp parallel for schedule(static) if (n > 65536)
    for (int y = 0; y < ny; ++y) {
        for (int x = 0; x < nx; ++x) {
            size_t i = IDX(y, x);
            double acc = adjoint ? g->dt2 * lap_at(g, work, y, x)
                                 : g->dt2 * g->c2[i] * lap_at(g, u, y, x);
            un[i] = 2.0 * u[i] - uo[i] + acc;
        }
    }
}

static void mur(const grid_t *g, const double *u, double *un)
{
    const int ny = g->ny, nx = g->nx;
    for (int y = 1; y < ny - 1; ++y) {
        un[IDX(y, 0)] = u[IDX(y, 1)] + g->k_left[y] * (un[IDX(y, 1)] - u[IDX(y, 0)]);
        un[IDX(y, nx - 1)] =
            u[IDX(y, nx - 2)] + g->k_right[y] * (un[IDX(y, nx - 2)] - u[IDX(y, nx - 1)]);
    }
    for (int x = 1; x < nx - 1; ++x) {
        un[IDX(0, x)] = u[IDX(1, x)] + g->k_bottom[x] * (un[IDX(1, x)] - u[IDX(0, x)]);
        un[IDX(ny - 1, x)] =
            u[IDX(ny - 2, x)] + g->k_top[x] * (un[IDX(ny - 2, x)] - u[IDX(ny - 1, x)]);
    }
    const int rr[4] = {0, 0, ny - 1, ny - 1}, cc[4] = {0, nx - 1, 0, nx - 1};
    const int ri[4] = {1, 1, ny - 2, ny - 2}, ci[4] = {1, nx - 2, 1, nx - 2};
    double v[4];
    for (int i = 0; i < 4; ++i) {
        double uc = u[IDX(rr[i], cc[i])];
        double fx = u[IDX(rr[i], ci[i])] + g->kx[i] * (un[IDX(rr[i], ci[i])] - uc);
        double fy = u[IDX(ri[i], cc[i])] + g->ky[i] * (un[IDX(ri[i], cc[i])] - uc);
        v[i] = 0.5 * (fx + fy);
    }
    for (int i = 0; i < 4; ++i) un[IDX(rr[i], cc[i])] = v[i];
}

static inline double gather(int nx, const double *f, int r0, int c0, const double *w, int R,
                            int r)
{
    return f[IDX(r0, c0)] * w[r] + f[IDX(r0, c0 + 1)] * w[R + r] +
           f[IDX(r0 + 1, c0)] * w[2 * R + r] + f[IDX(r0 + 1, c0 + 1)] * w[3 * R + r];
}

int oracle_n_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* Zero-padded Laplacian of one (ny, nx) field. */
void oracle_laplacian(int ny, int nx, double dx, double dy, const double *f, double *out)
{
    grid_t g;
    memset(&g, 0, sizeof g);
    g.ny = ny;
    g.nx = nx;
    g.idx2 = 1.0 / (dx * dx);
    g.idy2 = 1.0 / (dy * dy);
    g.cen = -2.0 * (g.idx2 + g.idy2);
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) out[IDX(y, x)] = lap_at(&g, f, y, x);
}

/* One step of the forward (adjoint = 0) or adjoint (adjoint = 1) operator with
 * an optional dense source added before the boundary, for one shot.
 * point_row < 0 disables the point source. */
int oracle_step(int ny, int nx, double dx, double dy, double dt, const double *c,
                const double *u_old, const double *u, double *u_next, int adjoint,
                int point_row, int point_col, double point_value, const double *dense)
{
    const size_t n = (size_t)ny * nx;
    double *c2 = malloc(n * sizeof(double));
    double *ks = malloc((2 * (size_t)ny + 2 * (size_t)nx) * sizeof(double));
    double *work = malloc(n * sizeof(double));
    if (!c2 || !ks || !work) {
        free(c2);
        free(ks);
        free(work);
        return 1;
    }
    grid_t g;
    grid_init(&g, ny, nx, dx, dy, dt, c, c2, ks);
    leapfrog(&g, u_old, u, u_next, adjoint, work);
    if (point_row >= 0) u_next[IDX(point_row, point_col)] += point_value;
    if (dense)
        for (size_t i = 0; i < n; ++i) u_next[i] += dense[i];
    mur(&g, u, u_next);
    free(c2);
    free(ks);
    free(work);
    return 0;
}

/* The transmitter of shot s at step value `val`: a point source at
 * (sr, sc), or with src_w (S, 4) the bilinear source w * val on the nodes
 * (sr + dy, sc + dx), weight index 2 dy + dx (w00, w10, w01, w11). */
static inline void add_transmitter(int nx, double *un, int s, int sr, int sc, double val,
                                   const double *src_w)
{
    if (!src_w) {
        un[IDX(sr, sc)] += val;
        return;
    }
    for (int k = 0; k < 4; ++k) un[IDX(sr + (k >> 1), sc + (k & 1))] += val * src_w[4 * s + k];
}

/* Forward shots. traces: (S, R, N) double. wave (optional): (N, S, ny, nx)
 * double, the field after every step. src_w (optional, (S, 4)): bilinear
 * transmitters with footprint corners (src_rows, src_cols). */
int oracle_forward(int ny, int nx, double dx, double dy, double dt, const double *c,
                   int n_shots, const int *src_rows, const int *src_cols,
                   const double *src_values, int n_steps, int n_rec, const int *row0,
                   const int *col0, const double *w, double *traces, double *wave,
                   const double *src_w)
{
    const size_t n = (size_t)ny * nx;
    double *c2 = malloc(n * sizeof(double));
    double *ks = malloc((2 * (size_t)ny + 2 * (size_t)nx) * sizeof(double));
    double *buf = calloc(3 * n, sizeof(double));
    if (!c2 || !ks || !buf) {
        free(c2);
        free(ks);
        free(buf);
        return 1;
    }
    grid_t g;
    grid_init(&g, ny, nx, dx, dy, dt, c, c2, ks);
    for (int s = 0; s < n_shots; ++s) {
        double *uo = buf, *u = buf + n, *un = buf + 2 * n;
        memset(buf, 0, 3 * n * sizeof(double));
        for (int t = 0; t < n_steps; ++t) {
            leapfrog(&g, uo, u, un, 0, NULL);
            add_transmitter(nx, un, s, src_rows[s], src_cols[s], src_values[t], src_w);
            mur(&g, u, un);
            for (int r = 0; r < n_rec; ++r)
                traces[((size_t)s * n_rec + r) * n_steps + t] =
                    gather(nx, un, row0[r], col0[r], w, n_rec, r);
            if (wave) memcpy(wave + ((size_t)t * n_shots + s) * n, un, n * sizeof(double));
            double *tmp = uo;
            uo = u;
            u = un;
            un = tmp;
        }
    }
    free(c2);
    free(ks);
    free(buf);
    return 0;
}

/* Misfit and adjoint-state gradient with respect to c^2, summed over shots.
 * observed: (S, R, N) double, tx[s]: muted receiver of shot s.
 * store_every K >= 1: imaging on steps with (n-1) % K == 0, weight K dt^2.
 * Outputs: *J, g_c2 (ny, nx), optional traces (S, R, N). src_w: as for
 * oracle_forward. */
int oracle_gradient(int ny, int nx, double dx, double dy, double dt, const double *c,
                    int n_shots, const int *src_rows, const int *src_cols,
                    const double *src_values, int n_steps, int n_rec, const int *row0,
                    const int *col0, const double *w, const double *observed,
                    const int *tx, int store_every, double *J_out, double *g_c2,
                    double *traces_out, const double *src_w)
{
    const size_t n = (size_t)ny * nx;
    double *c2 = malloc(n * sizeof(double));
    double *ks = malloc((2 * (size_t)ny + 2 * (size_t)nx) * sizeof(double));
    double *buf = calloc(5 * n, sizeof(double));
    double *hist = malloc((size_t)n_steps * n * sizeof(double));
    double *tr = malloc((size_t)n_rec * n_steps * sizeof(double));
    double *res = malloc((size_t)n_rec * n_steps * sizeof(double));
    if (!c2 || !ks || !buf || !hist || !tr || !res) {
        free(c2);
        free(ks);
        free(buf);
        free(hist);
        free(tr);
        free(res);
        return 1;
    }
    grid_t g;
    grid_init(&g, ny, nx, dx, dy, dt, c, c2, ks);
    memset(g_c2, 0, n * sizeof(double));
    double J = 0.0;
    const double scale = (double)store_every * g.dt2;
    for (int s = 0; s < n_shots; ++s) {
        double *uo = buf, *u = buf + n, *un = buf + 2 * n, *work = buf + 3 * n,
               *lap = buf + 4 * n;
        memset(buf, 0, 3 * n * sizeof(double));
        for (int t = 0; t < n_steps; ++t) {
            leapfrog(&g, uo, u, un, 0, NULL);
            add_transmitter(nx, un, s, src_rows[s], src_cols[s], src_values[t], src_w);
            mur(&g, u, un);
            for (int r = 0; r < n_rec; ++r)
                tr[(size_t)r * n_steps + t] = gather(nx, un, row0[r], col0[r], w, n_rec, r);
            memcpy(hist + (size_t)t * n, un, n * sizeof(double));
            double *tmp = uo;
            uo = u;
            u = un;
            un = tmp;
        }
        for (int r = 0; r < n_rec; ++r)
            for (int t = 0; t < n_steps; ++t) {
                size_t i = (size_t)r * n_steps + t;
                double d = tr[i] - observed[((size_t)s * n_rec + r) * n_steps + t];
                if (r == tx[s]) d = 0.0;
                res[i] = d;
                J += 0.5 * d * d;
            }
        if (traces_out) memcpy(traces_out + (size_t)s * n_rec * n_steps, tr,
                               (size_t)n_rec * n_steps * sizeof(double));
        /* adjoint, reverse time */
        uo = buf;
        u = buf + n;
        un = buf + 2 * n;
        memset(buf, 0, 3 * n * sizeof(double));
        for (int t = n_steps - 1; t >= 0; --t) {
            leapfrog(&g, uo, u, un, 1, work);
            for (int k = 0; k < 4; ++k) {
                const int dr = k >> 1, dc = k & 1;
                for (int r = 0; r < n_rec; ++r)
                    un[IDX(row0[r] + dr, col0[r] + dc)] +=
                        res[(size_t)r * n_steps + t] * w[(size_t)k * n_rec + r];
            }
            mur(&g, u, un);
            if (t > 0 && (t - 1) % store_every == 0) {
                const double *f = hist + (size_t)(t - 1) * n;
#pragma omp parallel for schedule(static) if (n > 65536)
                for (int y = 0; y < ny; ++y)
                    for (int x = 0; x < nx; ++x) lap[IDX(y, x)] = lap_at(&g, f, y, x);
                for (size_t i = 0; i < n; ++i) g_c2[i] += scale * un[i] * lap[i];
            }
            double *tmp = uo;
            uo = u;
            u = un;
            un = tmp;
        }
    }
    *J_out = J;
    free(c2);
    free(ks);
    free(buf);
    free(hist);
    free(tr);
    free(res);
    return 0;
}
