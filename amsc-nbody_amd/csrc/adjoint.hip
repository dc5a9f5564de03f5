// adjoint.hip -- energy of a simulation state and the reverse mode of the
// Barnes-Hut leapfrog.
//
// Energy.  Kinetic 1/2 sum m v^2 plus potential 1/2 sum m phi, with phi from
// an FP64 walk over the current tree (the same interaction lists as the
// forces of that tree), reduced per block in a fixed order and summed on the
// host in block order.
//
// Adjoint.  K kick-drift-kick steps from (x_0, v_0) (src/simulation.cpp:
// 100-119 and gpu/cuda_kernels/integrator.cu:51-106):
//   v_{j+1/2} = v_j + dt/2 a(x_j),  x_{j+1} = x_j + dt v_{j+1/2},
//   v_{j+1}   = v_{j+1/2} + dt/2 a(x_{j+1}).
// The tree of each force evaluation is held fixed (the forces are
// differentiated along their interaction lists).  With cotangents (xb, vb)
// of (x_K, v_K), the reverse sweep over j = K .. 1 is
//   ab = c_j vb  (c_K = dt/2, otherwise dt: a(x_j) feeds two half kicks),
//   xb += (da/dx)^T(x_j) ab,   vb += dt xb,
// and finally xb += (da/dx)^T(x_0) (dt/2 vb).  Forward stores (x, v) every
// `checkpoint_every` steps; the reverse sweep recomputes each segment from
// its checkpoint, keeping the segment's positions, and rebuilds the tree at
// each x_j for its vector-Jacobian product.
#include "sim_internal.hpp"

#include <cmath>
#include <vector>

namespace nbh {

void launch_potential_fp64(hipStream_t s, const Walk64Rec* w64,
                           const double4* wcenter, const int* d_wcount,
                           const double4* rec, const uint32_t* idx, int n,
                           double eps2, double G, bool contains, double* phi,
                           int* err);
void launch_energy_partials(hipStream_t s, const double* m, const double* vx,
                            const double* vy, const double* vz,
                            const double* phi, size_t n, double* partials,
                            int blocks);
void launch_force_vjp_fp64(hipStream_t s, const nbh_octree_view& o,
                           int max_level, const int2* pos,
                           const Walk64Rec* w64, const double4* wcenter,
                           const int* d_wcount, int wcount_host,
                           const double4* rec, const uint32_t* idx, int n,
                           double eps2, double G, bool contains,
                           const double* abx, const double* aby,
                           const double* abz, double* scratch, double* gx,
                           double* gy, double* gz, double* gm, int* split,
                           uint32_t* perm, uint32_t* pcost, int* err);
void launch_force_vjp_mixed(hipStream_t s, const nbh_octree_view& o, int max_level,
                            const int2* pos, const int4* aux, int4* wrec,
                            const int4* wcenter, const int* d_wcount, int* split,
                            int wcount_host, const int4* rec32, const double4* rec64,
                            const uint32_t* idx, const uint32_t* order, int n, float eps2,
                            double gunit, double munit, bool contains, bool soft,
                            const double* abx, const double* aby, const double* abz,
                            double* scratch, double* gx, double* gy, double* gz,
                            double* gm, uint32_t* perm, const uint32_t* cost, int* err);
size_t force_vjp_scratch_doubles(size_t n, size_t wcount, size_t n_nodes);

namespace {

// y = a * x (three planar arrays of n)
__global__ __launch_bounds__(256) void scale3_kernel(double a,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ y,
                                                     size_t n3)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n3)
        y[i] = a * x[i];
}

// y += a * x
__global__ __launch_bounds__(256) void axpy3_kernel(double a,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ y,
                                                    size_t n3)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n3)
        y[i] = fma(a, x[i], y[i]);
}

void scale3(hipStream_t s, double a, const double* x, double* y, size_t n)
{
    scale3_kernel<<<blocks_for(3 * n, 256), 256, 0, s>>>(a, x, y, 3 * n);
    NBH_HIP(hipGetLastError());
}

void axpy3(hipStream_t s, double a, const double* x, double* y, size_t n)
{
    axpy3_kernel<<<blocks_for(3 * n, 256), 256, 0, s>>>(a, x, y, 3 * n);
    NBH_HIP(hipGetLastError());
}

void copy(hipStream_t s, double* dst, const double* src, size_t n)
{
    NBH_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, s));
}

void require_fp64_walk(const nbh_sim* s, const char* what)
{
    if (mixed(s))
        fail(NBH_LOGIC_ERROR, std::string(what) +
                                  " needs an FP64 simulation (precision NBH_PRECISION_FP64)");
    if (s->state != 1)
        fail(NBH_LOGIC_ERROR, "Simulation must be initialized");
}

// Rebuilds the tree at the simulation's current positions (no kick/drift)
// and checks the build's device errors.
void rebuild_here(nbh_sim* s)
{
    do_rebuild(s, false);
    s->mark(PH_TRAVERSAL);
    s->mark(PH_KICK);
    s->mark(PH_COUNT);
    finish(s);
}

// xbar += (da/dx)^T abar and, when mbar is given, mbar += (da/dm)^T abar at
// the simulation's current tree.  abar and xbar are planar (x | y | z)
// arrays of n, mbar an array of n, original order.
void force_vjp(nbh_sim* s, const double* abar, double* xbar, DevBuf<double>& scratch,
               double* mbar = nullptr)
{
    const size_t n = s->n;
    const size_t M = static_cast<size_t>(s->nnodes);
    scratch.ensure(force_vjp_scratch_doubles(n, M, M));
    const bool contains = s->cfg.theta >= kContainsTheta;
    const double eps = s->cfg.softening;
    sim_fill_views(s); // the push-down reads node masses
    nbh_octree_view o = s->octree_view();
    // and every node's walk position, leaves included
    launch_leaf_positions(s->stream, o, s->poffset.p, s->pcount.p, s->ngroups,
                          s->level_arrays());
    if (mixed(s)) {
        // the mixed records' units: fixed-point steps per length, m_max
        const double scale = 2147483648.0 / s->rb->box[3];
        const double G = s->cfg.gravitational_constant;
        launch_force_vjp_mixed(
            s->stream, o, s->cfg.bits_per_axis, s->pos.p, s->aux.p, s->w32node.p,
            s->w32center.p, s->scalars.p + 3,
            walk_mixed_splits(static_cast<int>(n)) ? s->wsplit.p : nullptr,
            static_cast<int>(M), s->rec32.p, s->rec64.p, s->idx.p, s->horder.p,
            static_cast<int>(n), static_cast<float>(mixed_eps2(eps, scale)),
            G * s->rb->box[4] * scale * scale * scale, G * scale * scale, contains, eps > 0.0,
            abar, abar + n, abar + 2 * n, scratch.p, xbar, xbar + n, xbar + 2 * n, mbar,
            s->psched.p, s->psched.p + s->psched.n / 2, s->errors + ERR_TRAVERSAL);
    } else {
        launch_force_vjp_fp64(s->stream, o, s->cfg.bits_per_axis, s->pos.p,
                              s->w64.p, s->w64center.p, s->scalars.p + 3, static_cast<int>(M),
                              s->rec64.p, s->idx.p, static_cast<int>(n), eps * eps,
                              s->cfg.gravitational_constant, contains, abar,
                              abar + n, abar + 2 * n, scratch.p, xbar, xbar + n,
                              xbar + 2 * n, mbar, s->wsplit.p, s->psched.p,
                              s->psched.p ? s->psched.p + s->psched.n / 2 : nullptr,
                              s->errors + ERR_TRAVERSAL);
    }
    s->mark(PH_KICK);
    s->mark(PH_COUNT);
    finish(s);
}

void set_positions(nbh_sim* s, const double* x, const double* y, const double* z)
{
    copy(s->stream, s->x.p, x, s->n);
    copy(s->stream, s->y.p, y, s->n);
    copy(s->stream, s->z.p, z, s->n);
}

} // namespace
} // namespace nbh

struct nbh_adjoint {
    nbh_sim* sim = nullptr;
    size_t n = 0;
    int steps = 0;
    int every = 1;
    int segments = 0;
    bool forward_done = false;
    // Count of forward calls: the checkpoints, masses and tree a backward
    // replays belong to the most recent forward only.
    uint64_t generation = 0;
    nbh::DevBuf<double> ck;   // segments x 6n: x | y | z | vx | vy | vz
    nbh::DevBuf<double> seg;  // (every + 1) x 3n positions of one segment
    nbh::DevBuf<double> xb, vb, ab; // 3n each
    nbh::DevBuf<double> mb;         // n
    nbh::DevBuf<double> scratch;

    ~nbh_adjoint() { delete sim; }

    double* checkpoint(int k) { return ck.p + static_cast<size_t>(k) * 6 * n; }
    double* positions(int j) { return seg.p + static_cast<size_t>(j) * 3 * n; }

    void save_state(double* dst)
    {
        const hipStream_t st = sim->stream;
        double* src[6] = {sim->x.p, sim->y.p, sim->z.p, sim->vx.p, sim->vy.p, sim->vz.p};
        for (int k = 0; k < 6; ++k)
            nbh::copy(st, dst + k * n, src[k], n);
    }
    void load_state(const double* src)
    {
        const hipStream_t st = sim->stream;
        double* dst[6] = {sim->x.p, sim->y.p, sim->z.p, sim->vx.p, sim->vy.p, sim->vz.p};
        for (int k = 0; k < 6; ++k)
            nbh::copy(st, dst[k], src + k * n, n);
    }
    void save_positions(double* dst)
    {
        nbh::copy(sim->stream, dst, sim->x.p, n);
        nbh::copy(sim->stream, dst + n, sim->y.p, n);
        nbh::copy(sim->stream, dst + 2 * n, sim->z.p, n);
    }
    // Fresh start from the state in the simulation arrays: a(x) of the
    // current positions into the current buffer.
    void restart()
    {
        sim->state = 0;
        sim->steps = 0;
        sim->time = 0.0;
        nbh::sim_initialize(sim);
    }
};

using namespace nbh;

extern "C" {

int nbh_sim_energy(nbh_sim* s, double* kinetic, double* potential)
{
    return guard([&] {
        if (!s || !kinetic || !potential)
            fail(NBH_INVALID_ARGUMENT, "arguments must not be null");
        require_fp64_walk(s, "The energy");
        const size_t n = s->n;
        const int blocks = 256;
        DevBuf<double> phi, partials;
        phi.alloc(n);
        partials.alloc(2 * blocks);
        const bool contains = s->cfg.theta >= kContainsTheta;
        const double eps = s->cfg.softening;
        launch_potential_fp64(s->stream, s->w64.p, s->w64center.p, s->scalars.p + 3,
                              s->rec64.p,
                              s->idx.p, static_cast<int>(n), eps * eps,
                              s->cfg.gravitational_constant, contains, phi.p,
                              s->errors + ERR_TRAVERSAL);
        launch_energy_partials(s->stream, s->m.p, s->vx.p, s->vy.p, s->vz.p, phi.p,
                               n, partials.p, blocks);
        std::vector<double> h(2 * blocks);
        NBH_HIP(hipMemcpyAsync(h.data(), partials.p, h.size() * sizeof(double),
                               hipMemcpyDeviceToHost, s->stream));
        finish(s);
        double k = 0.0, p = 0.0;
        for (int b = 0; b < blocks; ++b) {
            k += h[2 * b];
            p += h[2 * b + 1];
        }
        *kinetic = k;
        *potential = p;
    });
}

int nbh_sim_potential(nbh_sim* s, double* d_phi)
{
    return guard([&] {
        if (!s || !d_phi)
            fail(NBH_INVALID_ARGUMENT, "arguments must not be null");
        require_fp64_walk(s, "The potential");
        const bool contains = s->cfg.theta >= kContainsTheta;
        const double eps = s->cfg.softening;
        launch_potential_fp64(s->stream, s->w64.p, s->w64center.p, s->scalars.p + 3,
                              s->rec64.p,
                              s->idx.p, static_cast<int>(s->n), eps * eps,
                              s->cfg.gravitational_constant, contains, d_phi,
                              s->errors + ERR_TRAVERSAL);
        finish(s);
    });
}

int nbh_sim_force_vjp(nbh_sim* s, const double* d_abx, const double* d_aby,
                      const double* d_abz, double* d_gx, double* d_gy,
                      double* d_gz, double* d_gm)
{
    return guard([&] {
        if (!s || !d_abx || !d_aby || !d_abz || !d_gx || !d_gy || !d_gz)
            fail(NBH_INVALID_ARGUMENT, "arguments must not be null");
        if (s->state != 1)
            fail(NBH_LOGIC_ERROR, "Simulation must be initialized");
        const size_t n = s->n;
        DevBuf<double> ab, g, scratch;
        ab.alloc(3 * n);
        g.alloc(4 * n);
        copy(s->stream, ab.p, d_abx, n);
        copy(s->stream, ab.p + n, d_aby, n);
        copy(s->stream, ab.p + 2 * n, d_abz, n);
        launch_fill_f64(s->stream, g.p, 0.0, 4 * n);
        force_vjp(s, ab.p, g.p, scratch, d_gm ? g.p + 3 * n : nullptr);
        copy(s->stream, d_gx, g.p, n);
        copy(s->stream, d_gy, g.p + n, n);
        copy(s->stream, d_gz, g.p + 2 * n, n);
        if (d_gm)
            copy(s->stream, d_gm, g.p + 3 * n, n);
        NBH_HIP(hipStreamSynchronize(s->stream));
    });
}

int nbh_adjoint_create(nbh_adjoint** out, const nbh_sim_config* config,
                       size_t count, int steps, int checkpoint_every,
                       void* hip_stream)
{
    return guard([&] {
        if (!out)
            fail(NBH_INVALID_ARGUMENT, "output pointer must not be null");
        if (!config)
            fail(NBH_INVALID_ARGUMENT, "config must not be null");
        if (config->precision != NBH_PRECISION_FP64 &&
            config->precision != NBH_PRECISION_MIXED)
            fail(NBH_INVALID_ARGUMENT, "unknown precision");
        if (steps < 1)
            fail(NBH_INVALID_ARGUMENT, "the adjoint needs at least one step");
        if (checkpoint_every < 1)
            fail(NBH_INVALID_ARGUMENT, "checkpoint_every must be positive");
        auto* a = new nbh_adjoint();
        try {
            a->sim = sim_new(config, count, hip_stream);
            a->n = count;
            a->steps = steps;
            a->every = checkpoint_every < steps ? checkpoint_every : steps;
            a->segments = (steps + a->every - 1) / a->every;
            a->ck.alloc(static_cast<size_t>(a->segments) * 6 * count);
            a->seg.alloc(static_cast<size_t>(a->every + 1) * 3 * count);
            a->xb.alloc(3 * count);
            a->vb.alloc(3 * count);
            a->ab.alloc(3 * count);
            a->mb.alloc(count);
        } catch (...) {
            delete a;
            throw;
        }
        *out = a;
    });
}

int nbh_adjoint_destroy(nbh_adjoint* a)
{
    return guard([&] { delete a; });
}

int nbh_adjoint_forward(nbh_adjoint* a, const double* d_mass, const double* d_x,
                        const double* d_y, const double* d_z, const double* d_vx,
                        const double* d_vy, const double* d_vz, double* d_x_out,
                        double* d_y_out, double* d_z_out, double* d_vx_out,
                        double* d_vy_out, double* d_vz_out)
{
    return guard([&] {
        if (!a || !d_mass || !d_x || !d_y || !d_z || !d_vx || !d_vy || !d_vz ||
            !d_x_out || !d_y_out || !d_z_out || !d_vx_out || !d_vy_out || !d_vz_out)
            fail(NBH_INVALID_ARGUMENT, "arguments must not be null");
        nbh_sim* s = a->sim;
        const size_t n = a->n;
        const hipStream_t st = s->stream;
        a->forward_done = false;
        ++a->generation;
        copy(st, s->m.p, d_mass, n);
        copy(st, s->x.p, d_x, n);
        copy(st, s->y.p, d_y, n);
        copy(st, s->z.p, d_z, n);
        copy(st, s->vx.p, d_vx, n);
        copy(st, s->vy.p, d_vy, n);
        copy(st, s->vz.p, d_vz, n);
        a->restart();
        for (int k = 0; k < a->steps; ++k) {
            if (k % a->every == 0)
                a->save_state(a->checkpoint(k / a->every));
            sim_step(s);
        }
        copy(st, d_x_out, s->x.p, n);
        copy(st, d_y_out, s->y.p, n);
        copy(st, d_z_out, s->z.p, n);
        copy(st, d_vx_out, s->vx.p, n);
        copy(st, d_vy_out, s->vy.p, n);
        copy(st, d_vz_out, s->vz.p, n);
        NBH_HIP(hipStreamSynchronize(st));
        a->forward_done = true;
    });
}

int nbh_adjoint_generation(const nbh_adjoint* a, uint64_t* generation)
{
    return guard([&] {
        if (!a || !generation)
            fail(NBH_INVALID_ARGUMENT, "arguments must not be null");
        *generation = a->generation;
    });
}

int nbh_adjoint_backward(nbh_adjoint* a, const double* d_xbar, const double* d_ybar,
                         const double* d_zbar, const double* d_vxbar,
                         const double* d_vybar, const double* d_vzbar,
                         double* d_x0bar, double* d_y0bar, double* d_z0bar,
                         double* d_vx0bar, double* d_vy0bar, double* d_vz0bar,
                         double* d_m0bar)
{
    return guard([&] {
        if (!a || !d_xbar || !d_ybar || !d_zbar || !d_vxbar || !d_vybar ||
            !d_vzbar || !d_x0bar || !d_y0bar || !d_z0bar || !d_vx0bar ||
            !d_vy0bar || !d_vz0bar)
            fail(NBH_INVALID_ARGUMENT, "arguments must not be null");
        if (!a->forward_done)
            fail(NBH_LOGIC_ERROR, "nbh_adjoint_backward needs a completed forward");
        nbh_sim* s = a->sim;
        const size_t n = a->n;
        const hipStream_t st = s->stream;
        const double dt = s->cfg.time_step;
        const int K = a->steps;
        double* xb = a->xb.p;
        double* vb = a->vb.p;
        double* ab = a->ab.p;
        // every force evaluation a(x_j, m) adds its mass cotangent
        double* mb = d_m0bar ? a->mb.p : nullptr;
        if (mb)
            launch_fill_f64(st, mb, 0.0, n);
        copy(st, xb, d_xbar, n);
        copy(st, xb + n, d_ybar, n);
        copy(st, xb + 2 * n, d_zbar, n);
        copy(st, vb, d_vxbar, n);
        copy(st, vb + n, d_vybar, n);
        copy(st, vb + 2 * n, d_vzbar, n);
        for (int sg = a->segments - 1; sg >= 0; --sg) {
            const int k0 = sg * a->every;
            const int k1 = (k0 + a->every < K) ? k0 + a->every : K;
            // recompute the segment, keeping x_{k0} .. x_{k1}
            a->load_state(a->checkpoint(sg));
            a->restart();
            a->save_positions(a->positions(0));
            for (int j = 1; j <= k1 - k0; ++j) {
                sim_step(s);
                a->save_positions(a->positions(j));
            }
            for (int j = k1; j > k0; --j) {
                scale3(st, (j == K) ? 0.5 * dt : dt, vb, ab, n);
                if (j != k1) {
                    const double* p = a->positions(j - k0);
                    set_positions(s, p, p + n, p + 2 * n);
                    rebuild_here(s);
                }
                force_vjp(s, ab, xb, a->scratch, mb);
                axpy3(st, dt, xb, vb, n);
            }
        }
        // a(x_0) feeds only the first half kick
        scale3(st, 0.5 * dt, vb, ab, n);
        const double* p0 = a->checkpoint(0);
        set_positions(s, p0, p0 + n, p0 + 2 * n);
        rebuild_here(s);
        force_vjp(s, ab, xb, a->scratch, mb);
        copy(st, d_x0bar, xb, n);
        copy(st, d_y0bar, xb + n, n);
        copy(st, d_z0bar, xb + 2 * n, n);
        copy(st, d_vx0bar, vb, n);
        copy(st, d_vy0bar, vb + n, n);
        copy(st, d_vz0bar, vb + 2 * n, n);
        if (mb)
            copy(st, d_m0bar, mb, n);
        NBH_HIP(hipStreamSynchronize(st));
    });
}

} // extern "C"
