// hv_adjoint_impl.h -- gradient of a dipole-moment misfit with respect to the
// Skyrme coefficients, by the discrete adjoint of the FP64 step loop
// (included at the end of hv_abi.hip: it uses that file's context and
// helpers).  No counterpart in the reference.
//
// Forward: hv_run_init + nsteps steps from the particles the context holds,
// with a checkpoint (x, k, U, phi) every `checkpoint_every` steps; the loss is
// L = sum_{s < nsteps} (D(t_s) - D_target(s))^2 with D = N Z / A (<z>_p - <z>_n)
// of the positions at the start of step s (scripts/programs.py:99).
//
// Reverse, per step s from the last to the first, with X and K the cotangents
// of the positions and momenta (a = dt / (2 hbar c), b = dt hbar c / mc^2):
//   K <- b X + K                       (drift, then the half kicks share K)
//   cot = a K (2 a K when step s - 1 also kicked with F(x_s))
//   (dx, dtheta) = VJP of F at x_s with cot (hv_adjoint.h)
//   X <- X + dx + dL/dx_s,  theta_bar += dtheta
// Each segment between checkpoints is replayed forward first (bitwise the
// same trajectory: the replay starts from the checkpointed warm potential) to
// store its positions and fields.  The adjoint Poisson solve treats the SOR as
// an exact solve of (-L_h) phi = 4 pi e^2 rho_p.

#include "hv_adjoint.h"

namespace {

constexpr int kAdjMaxIt = 2000;
constexpr int kAdjBlocks = 1024;

struct AdjBuf {
    double *X = nullptr, *Kb = nullptr, *cot = nullptr;  // 3N each
    unsigned long long *fbar = nullptr;                  // 3 x 2G, fixed point (k_adj_deposit_vec)
    double *fscale = nullptr;                            // its scale 2^k and 2^-k
    double *Ubar = nullptr, *rhobar = nullptr, *cbar = nullptr;  // 2G
    double *psi = nullptr, *psrc = nullptr;              // 2 Gh (split)
    double *partials = nullptr, *theta = nullptr, *amax = nullptr;
    double *seg_x = nullptr, *seg_U = nullptr, *seg_rho = nullptr;  // per replayed step
    double *ck = nullptr;                                // checkpoints
    double *init = nullptr;                              // the initial particles (9N)
    std::vector<void *> all;
};

int adj_alloc(hv_ctx *c, AdjBuf &b, void **p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        return set_err(c, HV_ERR_ALLOC, "adjoint: device allocation of %zu bytes failed", bytes);
    }
    b.all.push_back(*p);
    return HV_OK;
}

void adj_free(AdjBuf &b) {
    for (void *p : b.all) (void)hipFree(p);
    b.all.clear();
}

// one checkpoint: x, y, z, kx, ky, kz (N each), U (2G), smeared density (2G), phi (2 Gh)
size_t adj_ck_words(const hv_ctx *c) { return 6 * (size_t)c->total + 4 * (size_t)c->G + 2 * (size_t)c->sp.Gh; }

int adj_checkpoint(hv_ctx *c, double *dst, bool save) {
    const size_t N = c->total;
    const hipMemcpyKind kind = hipMemcpyDeviceToDevice;
    for (int q = 0; q < 6; q++) {
        double *a = dst + q * N, *p = (double *)c->part[q];
        HV_HIP(c, hipMemcpyAsync(save ? a : p, save ? p : a, sizeof(double) * N, kind, c->stream));
    }
    double *u = dst + 6 * N, *rh = u + 2 * c->G, *ph = rh + 2 * c->G;
    HV_HIP(c, hipMemcpyAsync(save ? u : c->U, save ? c->U : u, sizeof(double) * 2 * c->G, kind, c->stream));
    HV_HIP(c, hipMemcpyAsync(save ? rh : c->density, save ? c->density : rh, sizeof(double) * 2 * c->G, kind,
                             c->stream));
    HV_HIP(c, hipMemcpyAsync(save ? ph : c->phi, save ? c->phi : ph, sizeof(double) * 2 * c->sp.Gh, kind, c->stream));
    return HV_OK;
}

// VJP of the force map at the state (x, U, rho) for the cotangent b.cot:
// adds dF^T cot to b.X (and the loss term gz), and dtheta to b.theta.
int adj_force_vjp(hv_ctx *c, AdjBuf &b, const double *x, const double *U, const double *rho, const hv_skyrme *s,
                  double omega_adj, double gz_p, double gz_n) {
    const int N = c->total;
    const long long G = c->G, G2 = 2 * G;
    Geom<double> g = make_geom<double>(c);
    Particles<double> pp{};
    pp.x = (double *)x;
    pp.y = (double *)x + N;
    pp.z = (double *)x + 2LL * N;
    int rc = fill<unsigned long long>(c, b.fbar, 0ULL, 3 * G2);
    if (rc) return rc;
    // fixed-point scale from max |cot| (k_adj_fix_scale)
    k_adj_absmax<<<kAdjBlocks, kThreads, 0, c->stream>>>(b.cot, 3LL * N, b.amax);
    HV_CHECK_LAUNCH(c);
    k_adj_absmax<<<1, kThreads, 0, c->stream>>>(b.amax, kAdjBlocks, b.amax + kAdjBlocks);
    HV_CHECK_LAUNCH(c);
    k_adj_fix_scale<<<1, 1, 0, c->stream>>>(b.amax + kAdjBlocks, N, b.fscale);
    HV_CHECK_LAUNCH(c);
    k_adj_deposit_vec<<<blocks_for(N), kThreads, 0, c->stream>>>(pp, c->np, N, g, b.cot, b.fscale, b.fbar);
    HV_CHECK_LAUNCH(c);
    k_adj_fdm_t<<<blocks_for(G2), kThreads, 0, c->stream>>>(b.fbar, b.fscale, b.Ubar, g);
    HV_CHECK_LAUNCH(c);
    const unsigned nbs = blocks_for(G);
    k_adj_skyrme<<<nbs, kThreads, 0, c->stream>>>(rho, b.Ubar, b.rhobar, s->A, s->B, s->C, s->gamma, G, b.partials);
    HV_CHECK_LAUNCH(c);
    k_adj_theta_sum<<<1, 64, 0, c->stream>>>(b.partials, (int)nbs, b.theta);
    HV_CHECK_LAUNCH(c);
    // adjoint Poisson: psi = 4 pi e^2 (-L_h)^-1 Ubar_p (zero faces), warm-started
    k_rho_split<double><<<blocks_for(G), kThreads, 0, c->stream>>>(b.Ubar, b.psrc, G, c->sp);
    HV_CHECK_LAUNCH(c);
    k_adj_absmax<<<kAdjBlocks, kThreads, 0, c->stream>>>(b.Ubar, G, b.amax);
    HV_CHECK_LAUNCH(c);
    k_adj_absmax<<<1, kThreads, 0, c->stream>>>(b.amax, kAdjBlocks, b.amax + kAdjBlocks);
    HV_CHECK_LAUNCH(c);
    double umax = 0.0;
    HV_HIP(c, hipMemcpyAsync(&umax, b.amax + kAdjBlocks, sizeof umax, hipMemcpyDeviceToHost, c->stream));
    HV_HIP(c, hipStreamSynchronize(c->stream));
    if (umax > 0.0) {
        SorCoef cf = make_sor_coef(c, omega_adj, 1.0);
        // stop once the updates are 1e-10 of the first sweep's scale
        cf.tol = 1e-10 * omega_adj * cf.src * umax * cf.inv_dx2;
        const SorTiles tl = make_sor_tiles(c);
        const dim3 sblk(64, kSorRows, 1), sgrd((unsigned)(tl.per_xcd * kXcds), 1, 1);
        int iters = 0;
        double md = 0.0;
        rc = sor_iterate(c, cf, kAdjMaxIt, &iters, &md, [&](int it) {
            for (int colour = 0; colour < 2; colour++)
                k_sor_colour<double><<<sgrd, sblk, 0, c->stream>>>(b.psi, b.psrc, c->sp, tl, cf, colour, it, c->maxd,
                                                                   c->ctl);
        });
        if (rc) return rc;
        k_adj_add_split<<<blocks_for(G), kThreads, 0, c->stream>>>(b.psi, b.rhobar, G, c->sp);
        HV_CHECK_LAUNCH(c);
    }
    // smearing is symmetric: cbar = smear(rhobar)
    const SmearTab<double> *tb = (const SmearTab<double> *)c->smear_tab;
    const int nx = c->world.n[0], ny = c->world.n[1], nz = c->world.n[2];
    k_smear_fused<double, double><<<(unsigned)smear_blocks(c->stf), kSmearThreads, c->stf_lds, c->stream>>>(
        b.rhobar, b.cbar, tb, nx, ny, nz, c->stf);
    HV_CHECK_LAUNCH(c);
    k_adj_particle<<<blocks_for(N), kThreads, 0, c->stream>>>(pp, c->np, N, g, U, b.cot, b.cbar, b.X, gz_p, gz_n);
    HV_CHECK_LAUNCH(c);
    return HV_OK;
}

int adj_setup(hv_ctx *c, AdjBuf &b, int seg_len, int nck) {
    const size_t N = c->total, G = c->G, G2 = 2 * G, Gh2 = 2 * (size_t)c->sp.Gh;
    int rc = HV_OK;
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.X, sizeof(double) * 3 * N);
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.Kb, sizeof(double) * 3 * N);
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.cot, sizeof(double) * 3 * N);
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.fbar, sizeof(unsigned long long) * 3 * G2);
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.fscale, sizeof(double) * 2);
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.Ubar, sizeof(double) * G2);
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.rhobar, sizeof(double) * G2);
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.cbar, sizeof(double) * G2);
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.psi, sizeof(double) * Gh2);
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.psrc, sizeof(double) * Gh2);
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.partials, sizeof(double) * 4 * blocks_for((long long)G));
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.theta, sizeof(double) * 4);
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.amax, sizeof(double) * (kAdjBlocks + 1));
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.seg_x, sizeof(double) * 3 * N * (size_t)seg_len);
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.seg_U, sizeof(double) * G2 * (size_t)seg_len);
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.seg_rho, sizeof(double) * G2 * (size_t)seg_len);
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.ck, sizeof(double) * adj_ck_words(c) * (size_t)nck);
    rc = rc ? rc : adj_alloc(c, b, (void **)&b.init, sizeof(double) * 9 * N);
    if (rc) return rc;
    rc = fill<double>(c, b.X, 0.0, 3 * (long long)N);
    rc = rc ? rc : fill<double>(c, b.Kb, 0.0, 3 * (long long)N);
    rc = rc ? rc : fill<double>(c, b.psi, 0.0, (long long)Gh2);
    rc = rc ? rc : fill<double>(c, b.theta, 0.0, 4);
    return rc;
}

int dipole_loss_grad(hv_ctx *c, const hv_skyrme *s, const hv_params *p, int nsteps, const double *target,
                     int ck_every, double *loss, double *d_series, double *grad) {
    if (c->precision != HV_PREC_F64) return set_err(c, HV_ERR_STATE, "the adjoint runs on an FP64 context");
    if (nsteps <= 0 || ck_every <= 0 || !target || !loss)
        return set_err(c, HV_ERR_ARG, "need nsteps > 0, checkpoint_every > 0, target and loss");
    int rc = check_params(c, s, p);
    if (rc) return rc;
    const int N = c->total;
    const int nseg = (nsteps + ck_every - 1) / ck_every;
    // the particles in the caller's order, never re-sorted during the run
    rc = restore_order(c);
    if (rc) return rc;
    AdjBuf b;
    // on every return (the HV_HIP / HV_CHECK_LAUNCH early returns included):
    // restore the sort cadence, drain the stream, free the adjoint buffers
    struct Guard {
        hv_ctx *c;
        AdjBuf &b;
        int sort_saved;
        ~Guard() {
            c->sort_every = sort_saved;
            (void)hipStreamSynchronize(c->stream);
            adj_free(b);
        }
    } guard{c, b, c->sort_every};
    c->sort_every = 0;
    rc = ensure_smear<double>(c, p->sigma_r, p->part_per_nucleon);
    if (!rc) rc = ensure_maxd(c, kAdjMaxIt > p->sor_max_it ? kAdjMaxIt : p->sor_max_it);
    if (!rc) rc = ensure_pinned(c, (size_t)(kAdjMaxIt > p->sor_max_it ? kAdjMaxIt : p->sor_max_it) + 8);
    if (!rc) rc = adj_setup(c, b, ck_every, nseg);
    if (rc) return rc;
    auto done = [](int r) { return r; };
    for (int q = 0; q < 9; q++)
        HV_HIP(c, hipMemcpyAsync(b.init + (size_t)q * N, c->part[q], sizeof(double) * N, hipMemcpyDeviceToDevice,
                                 c->stream));
    // ---- forward, with checkpoints
    int warn = run_init<double>(c, s, p, nullptr);
    if (warn < 0) return done(warn);
    std::vector<double> stats((size_t)nsteps * 4);
    const size_t ckw = adj_ck_words(c);
    for (int j = 0; j < nseg; j++) {
        const int s0 = j * ck_every, len = nsteps - s0 < ck_every ? nsteps - s0 : ck_every;
        rc = adj_checkpoint(c, b.ck + ckw * j, true);
        if (rc) return done(rc);
        rc = run_steps<double>(c, s, p, s0, len, stats.data() + (size_t)s0 * 4, nullptr);
        if (rc < 0) return done(rc);
        warn |= rc;
    }
    const double dfac = (double)p->n * p->z / (p->n + p->z);
    std::vector<double> resid(nsteps);
    double L = 0.0;
    for (int t = 0; t < nsteps; t++) {
        const double d = dfac * (stats[4 * (size_t)t] - stats[4 * (size_t)t + 1]);
        if (d_series) d_series[t] = d;
        resid[t] = d - target[t];
        L += resid[t] * resid[t];
    }
    *loss = L;
    if (grad) {
        // ---- reverse sweep, segment by segment
        const double dt = p->t_f / p->steps;
        const double a = dt / (2.0 * HV_HBARC), bdr = (dt * HV_HBARC) / HV_MC2;
        const double omega_adj = 2.0 / (1.0 + sin(HV_PI / c->world.n[0]));
        const size_t N3 = 3 * (size_t)N, G2 = 2 * (size_t)c->G;
        for (int j = nseg - 1; j >= 0; j--) {
            const int s0 = j * ck_every, len = nsteps - s0 < ck_every ? nsteps - s0 : ck_every;
            rc = adj_checkpoint(c, b.ck + ckw * j, false);
            if (rc) return done(rc);
            c->phi_warm = true;
            // the boost fires once, at the first step with step dt >= t_exc: it
            // has fired before step s0 when step s0 - 1 already qualified
            c->excited = (s0 > 0 && (s0 - 1) * dt >= p->t_exc) ? 1 : 0;
            for (int q = 0; q < len; q++) {
                // state at the start of step s0 + q: x (3N), U, rho of x
                for (int a3 = 0; a3 < 3; a3++)
                    HV_HIP(c, hipMemcpyAsync(b.seg_x + N3 * q + (size_t)a3 * N, c->part[a3], sizeof(double) * N,
                                             hipMemcpyDeviceToDevice, c->stream));
                HV_HIP(c, hipMemcpyAsync(b.seg_U + G2 * q, c->U, sizeof(double) * G2, hipMemcpyDeviceToDevice,
                                         c->stream));
                HV_HIP(c, hipMemcpyAsync(b.seg_rho + G2 * q, c->density, sizeof(double) * G2,
                                         hipMemcpyDeviceToDevice, c->stream));
                if (q + 1 < len) {
                    rc = run_steps<double>(c, s, p, s0 + q, 1, nullptr, nullptr);
                    if (rc < 0) return done(rc);
                }
            }
            for (int q = len - 1; q >= 0; q--) {
                const int st = s0 + q;
                const double am = a * (st >= 1 ? 2.0 : 1.0);
                k_adj_kbar<<<blocks_for((long long)N3), kThreads, 0, c->stream>>>(b.Kb, b.X, b.cot, bdr, am,
                                                                                  (long long)N3);
                HV_CHECK_LAUNCH(c);
                const double gz = 2.0 * resid[st] * dfac;
                rc = adj_force_vjp(c, b, b.seg_x + N3 * q, b.seg_U + G2 * q, b.seg_rho + G2 * q, s, omega_adj,
                                   gz / c->np, -gz / c->nn);
                if (rc) return done(rc);
            }
        }
        double th[4];
        HV_HIP(c, hipMemcpyAsync(th, b.theta, sizeof th, hipMemcpyDeviceToHost, c->stream));
        HV_HIP(c, hipStreamSynchronize(c->stream));
        for (int k = 0; k < 4; k++) grad[k] = th[k];
    }
    // leave the initial particles in place for the next call
    for (int q = 0; q < 9; q++)
        HV_HIP(c, hipMemcpyAsync(c->part[q], b.init + (size_t)q * N, sizeof(double) * N, hipMemcpyDeviceToDevice,
                                 c->stream));
    return done(warn);
}

}  // namespace

extern "C" {

int hv_dipole_loss_grad(hv_ctx *c, const hv_skyrme *skyrme, const hv_params *params, int nsteps,
                        const double *d_target, int checkpoint_every, double *loss, double *d_series, double *grad) {
    if (!c) return HV_ERR_ARG;
    return dipole_loss_grad(c, skyrme, params, nsteps, d_target, checkpoint_every, loss, d_series, grad);
}

}  // extern "C"
