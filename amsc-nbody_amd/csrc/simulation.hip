// simulation.hip -- the Simulation object: the whole Barnes-Hut leapfrog
// step on one stream with two host synchronisations per step.
//
// Replaces Simulation (src/simulation.cpp:70-225) and its device state
// (gpu/cuda_kernels/simulation.cu:14-142).  The reference reallocates every
// staged structure and synchronises about twenty times per step; here every
// buffer is sized once for N (the octree grows on demand), group and node
// counts stay on the device, and the host reads them once, after the
// radix-to-octree plan, to size the octree launches.  The second
// synchronisation ends the step and reports any device error.
//
// Step order (src/simulation.cpp:100-119): kick-drift with the current
// accelerations (fused with the bounding box), rebuild (Morton keys, sort,
// groups, Karras, plan, octree, mass/COM, walk layout), forces into the next
// buffer, closing kick, swap.
#include "sim_internal.hpp"
#include "hilbert.hpp"

#include <cmath>
#include <cstring>
#include <limits>

namespace nbh {

namespace {

const char* kPhaseNames[PH_COUNT] = {
    "kick_drift_bbox", "encode", "sort",  "gather",     "groups",    "karras",
    "plan",            "octree", "com",   "walk_layout", "traversal", "kick"};

} // namespace

bool mixed(const nbh_sim* s) { return s->cfg.precision == NBH_PRECISION_MIXED; }

void alloc_state(nbh_sim* s)
{
    const size_t n = s->n;
    for (auto* b : {&s->m, &s->x, &s->y, &s->z, &s->vx, &s->vy, &s->vz})
        b->alloc(n);
    for (auto& a : s->acc)
        for (auto& c : a)
            c.alloc(n);
    s->keys.alloc(n * s->key_bytes);
    s->keys_alt.alloc(n * s->key_bytes);
    s->unique.alloc(n * s->key_bytes);
    s->idx.alloc(n);
    s->idx_alt.alloc(n);
    s->rec64.alloc(n);
    s->pos4.alloc(n);
    // both traversals take their packets along the Hilbert curve
    s->horder.alloc(n);
    const HilbertTables& tab = hilbert_tables();
    s->htab.alloc(sizeof(HilbertTables));
    NBH_HIP(hipMemcpyAsync(s->htab.p, &tab, sizeof(HilbertTables),
                           hipMemcpyHostToDevice, s->stream));
    if (mixed(s)) {
        s->rec32.alloc(n);
        if (walk_mixed_splits(static_cast<int>(n))) {
            s->wsplit.alloc(kWalkSplitInts);
            s->wpart.alloc((n + 63) / 64 * kWalkSplitMax * 64 * 3);
        }
    }
    if (mixed(s)) {
        // the traversal's units: packets, or packets times subtrees if split
        const size_t units = (n + 63) / 64 * (walk_mixed_splits(static_cast<int>(n))
                                                  ? kWalkSplitMax : 1);
        s->psched.alloc(2 * units);
        NBH_HIP(hipMemsetAsync(s->psched.p + units, 0, units * sizeof(uint32_t), s->stream));
    } else if (walk_mixed_splits(static_cast<int>(n))) {
        // the FP64 force VJP splits its units likewise (the force walk does not)
        const size_t units = (n + 63) / 64 * kWalkSplitMax;
        s->wsplit.alloc(kWalkSplitInts);
        s->psched.alloc(2 * units);
        NBH_HIP(hipMemsetAsync(s->psched.p + units, 0, units * sizeof(uint32_t), s->stream));
    }
    s->gfirst.alloc(n);
    s->gcount.alloc(n);
    s->left.alloc(n);
    s->right.alloc(n);
    s->rfirst.alloc(n);
    s->rlast.alloc(n);
    s->plen.alloc(n);
    s->tparent.alloc(2 * n);
    s->pplen.alloc(2 * n);
    s->plevel.alloc(2 * n);
    s->pcount.alloc(2 * n);
    s->poffset.alloc(2 * n);
    const size_t hist_w = 8 * 256;
    const size_t sort_w = sort_status_words(n, s->passes);
    const size_t grp_w = 2 * scan_status_words(n);
    const size_t plan_w = 2 * scan_status_words(2 * n);
    // The per-step cleared words end with the readback's error words; the
    // rest of the readback block follows them, outside the cleared range.
    const size_t rb_w = sizeof(Readback) / sizeof(uint32_t);
    size_t off = hist_w + sort_w;
    off = (off + 1) & ~size_t(1); // 64-bit words must be 8-byte aligned
    const size_t group_off = off;
    off += grp_w;
    const size_t plan_off = off;
    off += plan_w;
    s->zero.alloc(off + rb_w + 2);
    uint32_t* z = s->zero.p;
    s->hist = z;
    s->sort_status = z + hist_w;
    s->group_status = reinterpret_cast<unsigned long long*>(z + group_off);
    s->plan_status = reinterpret_cast<unsigned long long*>(z + plan_off);
    s->mid = reinterpret_cast<Readback*>(z + off);
    s->errors = s->mid->errors;
    s->box = s->mid->box;
    s->zero_words = off + kReadbackClearedWords;
    s->partials.alloc(bbox_partial_doubles());
    s->scalars.alloc(16);
    s->counters.alloc(4);
    s->walkq.alloc(kWalkQueueInts);
    NBH_HIP(hipMemsetAsync(s->scalars.p, 0, 16 * sizeof(int), s->stream));
    NBH_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->rb), sizeof(Readback),
                          hipHostMallocDefault));
    for (auto& e : s->ev)
        NBH_HIP(hipEventCreate(&e));
}

void ensure_octree(nbh_sim* s, size_t m)
{
    if (m <= s->mcap)
        return;
    const size_t cap = m + m / 4 + 1024;
    s->children.alloc(8 * cap);
    for (auto* b : {&s->oparent, &s->level, &s->ofirst, &s->ocount, &s->pending,
                    &s->child_count, &s->order})
        b->alloc(cap);
    s->sum4.alloc(cap);
    s->aux.alloc(cap);
    s->crit.alloc(cap);
    s->link.alloc(cap);
    s->pos.alloc(cap);
    s->prefix.alloc(cap * s->key_bytes);
    for (auto* b : {&s->mass, &s->comx, &s->comy, &s->comz, &s->cenx, &s->ceny,
                    &s->cenz, &s->half})
        b->alloc(cap);
    if (mixed(s)) {
        s->w32node.alloc(2 * cap); // interleaved {node, link} pairs
        if (s->cfg.theta >= kContainsTheta)
            s->w32center.alloc(cap);
    } else {
        s->w64.alloc(cap);
        if (s->cfg.theta >= kContainsTheta)
            s->w64center.alloc(cap);
    }
    if (cap > 0x7FFFFFFFull)
        fail(NBH_OVERFLOW_ERROR, "octree node count exceeds the walk's position range");
    launch_walk_sentinel(s->stream, mixed(s) ? s->w32node.p : nullptr,
                         mixed(s) ? nullptr : s->w64.p, static_cast<int>(cap));
    s->mcap = cap;
}

template <int BITS>
void rebuild(nbh_sim* s, bool kick_drift)
{
    using K = typename Layout<BITS>::Key;
    hipStream_t st = s->stream;
    const size_t n = s->n;
    const int ni = static_cast<int>(n);
    double* const* a = nullptr;
    (void)a;
    s->mark(PH_KICK_DRIFT_BBOX);
    s->views_fresh = false;
    launch_fill_u32(st, s->zero.p, 0u, s->zero_words);
    if (kick_drift) {
        auto& c = s->acc[s->cur];
        launch_kick_drift_bbox(st, s->x.p, s->y.p, s->z.p, s->vx.p, s->vy.p,
                               s->vz.p, c[0].p, c[1].p, c[2].p, n,
                               s->cfg.time_step, s->box, s->partials.p,
                               s->errors + ERR_BBOX, s->m.p, s->pos4.p);
    } else {
        nbh_context tmp;
        tmp.stream = st;
        launch_bbox(&tmp, s->x.p, s->y.p, s->z.p, n, s->box, s->partials.p,
                    s->errors + ERR_BBOX, s->m.p, s->pos4.p);
    }
    s->mark(PH_ENCODE);
    K* keys = reinterpret_cast<K*>(s->keys.p);
    launch_encode<BITS>(st, s->x.p, s->y.p, s->z.p, n, s->box, keys,
                        s->idx.p, s->hist);
    s->mark(PH_SORT);
    launch_radix_passes<K>(st, keys, s->idx.p, reinterpret_cast<K*>(s->keys_alt.p),
                           s->idx_alt.p, n, s->passes, s->hist, s->sort_status);
    s->mark(PH_GATHER);
    launch_gather(st, s->idx.p, s->m.p, s->x.p, s->y.p, s->z.p, n, s->rec64.p,
                  mixed(s) ? s->rec32.p : nullptr, s->box, s->pos4.p);
    s->mark(PH_GROUPS);
    int* d_ng = &s->mid->ngroups;
    int* d_nm = &s->mid->nnodes;
    K* unique = reinterpret_cast<K*>(s->unique.p);
    launch_leaf_groups<K>(st, keys, ni, unique, s->gfirst.p, s->gcount.p, d_ng,
                          s->group_status);
    s->mark(PH_KARRAS);
    launch_karras<K>(st, unique, d_ng, ni, s->left.p, s->right.p, s->tparent.p,
                     s->rfirst.p, s->rlast.p, s->plen.p, s->pplen.p);
    s->mark(PH_PLAN);
    launch_plan_scan<BITS>(st, d_ng, ni, s->tparent.p, s->plen.p, unique,
                           s->plevel.p, s->pcount.p, s->poffset.p, d_nm,
                           s->errors + ERR_PLAN, s->plan_status, s->mid->level_count,
                           s->pplen.p);
    // The one mid-step readback: stage errors, group and node counts, the
    // internal nodes per level and the box, in one copy.
    NBH_HIP(hipMemcpyAsync(s->rb, s->mid, sizeof(Readback), hipMemcpyDeviceToHost, st));
    NBH_HIP(hipStreamSynchronize(st));
    if (s->rb->errors[ERR_BBOX] != 0)
        fail(NBH_DOMAIN_ERROR,
             "Particle positions must be finite to compute a bounding box.");
    if (s->rb->errors[ERR_PLAN] != 0)
        fail(NBH_RUNTIME_ERROR, stage_message(ERR_PLAN, s->rb->errors[ERR_PLAN]),
             s->rb->errors[ERR_PLAN]);
    s->ngroups = s->rb->ngroups;
    if (s->ngroups <= 0 || s->ngroups > ni)
        fail(NBH_RUNTIME_ERROR,
             "Leaf grouping produced an invalid group count (" +
                 std::to_string(s->ngroups) + " groups, " +
                 std::to_string(s->rb->nnodes) + " nodes for " +
                 std::to_string(ni) + " particles)");
    if (s->rb->nnodes <= 0)
        fail(NBH_OVERFLOW_ERROR,
             "The computed octree node count does not fit in int");
    s->nnodes = s->rb->nnodes;
    s->starts = level_starts(s->rb->level_count);
    ensure_octree(s, static_cast<size_t>(s->nnodes));
    s->mark(PH_OCTREE);
    nbh_octree_view o = s->octree_view();
    const LevelArrays la = s->level_arrays();
    launch_octree_build<BITS>(st, o, s->ngroups, s->tparent.p, s->rfirst.p,
                              unique, s->gfirst.p, s->gcount.p, s->plevel.p,
                              s->pcount.p, s->poffset.p, s->errors + ERR_OCTREE,
                              false);
    launch_level_scatter(st, d_ng, s->ngroups, s->plevel.p, s->pcount.p, s->poffset.p,
                         s->mid->level_count, s->mid->level_cursor, s->order.p);
    s->mark(PH_COM);
    const bool contains = s->cfg.theta >= kContainsTheta;
    launch_level_passes<BITS>(
        st, o, s->starts, mixed(s), s->rec64.p, s->rec32.p, s->cfg.theta, s->box, contains,
        s->cfg.softening, la, s->order.p, s->poffset.p, s->pcount.p, s->ngroups,
        s->cfg.gravitational_constant, s->w64.p, s->w64center.p, s->w32node.p,
        s->w32center.p, s->scalars.p + 3, s->timing ? s->ev[PH_WALK] : nullptr,
        s->horder.p, s->htab.p);
}

void forces(nbh_sim* s, int buffer, unsigned long long* counters)
{
    hipStream_t st = s->stream;
    auto& a = s->acc[buffer];
    const int ni = static_cast<int>(s->n);
    const bool contains = s->cfg.theta >= kContainsTheta;
    const double eps = s->cfg.softening;
    if (!counters)
        s->mark(PH_TRAVERSAL);
    if (mixed(s)) {
        // Records hold m / box[4] on the 2^31 grid of the box: the FP32 sums
        // times G box[4] scale^2 are the physical accelerations.
        const double scale = 2147483648.0 / s->rb->box[3];
        const double out_scale = s->cfg.gravitational_constant * s->rb->box[4] * scale * scale;
        launch_walk_mixed(st, s->w32node.p, s->mcap, s->w32center.p,
                          s->scalars.p + 3, s->rec32.p, s->idx.p, s->horder.p, ni,
                          static_cast<float>(mixed_eps2(eps, scale)), out_scale, contains,
                          eps > 0.0,
                          a[0].p, a[1].p, a[2].p, s->walkq.p, s->wsplit.p, s->wpart.p,
                          s->psched.p, s->psched.p + s->psched.n / 2,
                          s->errors + ERR_TRAVERSAL, counters);
    } else {
        launch_walk_fp64(st, s->w64.p, s->mcap, s->w64center.p, s->scalars.p + 3,
                         s->rec64.p, s->idx.p, s->horder.p, ni, eps * eps,
                         s->cfg.gravitational_constant, contains, s->rb->box[3], a[0].p,
                         a[1].p, a[2].p, s->walkq.p, s->errors + ERR_TRAVERSAL, counters);
    }
}

// Second synchronisation of a step: device errors of the octree, the
// mass/COM climb and the traversal.
void finish(nbh_sim* s)
{
    NBH_HIP(hipMemcpyAsync(s->rb->errors, s->errors,
                           ERR_WORD_COUNT * sizeof(int), hipMemcpyDeviceToHost,
                           s->stream));
    NBH_HIP(hipStreamSynchronize(s->stream));
    for (int w : {ERR_OCTREE, ERR_PHYSICS, ERR_TRAVERSAL}) {
        const int code = s->rb->errors[w];
        if (code != 0)
            fail(NBH_RUNTIME_ERROR, stage_message(w, code), code);
    }
    if (s->timing) {
        for (int p = 0; p < PH_COUNT; ++p) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, s->ev[p], s->ev[p + 1]) == hipSuccess)
                s->phase_ms[p] = ms;
            else
                s->phase_ms[p] = 0.0;
        }
    }
}

void sim_fill_views(nbh_sim* s)
{
    if (s->views_fresh || s->nnodes <= 0)
        return;
    nbh_octree_view o = s->octree_view();
    if (s->cfg.bits_per_axis == 10)
        launch_octree_views<10>(s->stream, o, s->level_arrays(), s->rec64.p, s->box);
    else
        launch_octree_views<21>(s->stream, o, s->level_arrays(), s->rec64.p, s->box);
    s->views_fresh = true;
}

void do_rebuild(nbh_sim* s, bool kick_drift)
{
    if (s->cfg.bits_per_axis == 10)
        rebuild<10>(s, kick_drift);
    else
        rebuild<21>(s, kick_drift);
}

void sim_initialize(nbh_sim* s)
{
    try {
        do_rebuild(s, false);
        forces(s, 1 - s->cur, nullptr);
        s->mark(PH_KICK);
        s->mark(PH_COUNT);
        finish(s);
        s->cur = 1 - s->cur;
        s->state = 1;
    } catch (...) {
        s->state = 2;
        throw;
    }
}

void sim_step(nbh_sim* s)
{
    try {
        do_rebuild(s, true);
        const int next = 1 - s->cur;
        forces(s, next, nullptr);
        s->mark(PH_KICK);
        auto& a = s->acc[next];
        launch_kick(s->stream, s->vx.p, s->vy.p, s->vz.p, a[0].p, a[1].p,
                    a[2].p, s->n, s->cfg.time_step);
        s->mark(PH_COUNT);
        finish(s);
        s->cur = next;
        s->time += s->cfg.time_step;
        ++s->steps;
    } catch (...) {
        s->state = 2;
        throw;
    }
}

nbh_sim* sim_new(const nbh_sim_config* config, size_t count, void* hip_stream)
{
    validate_config(config);
    if (count == 0)
        fail(NBH_INVALID_ARGUMENT, "Simulation requires at least one particle");
    if (count > static_cast<size_t>(std::numeric_limits<int>::max()) / 2 ||
        count >= (1u << 30))
        fail(NBH_OVERFLOW_ERROR, "Particle count exceeds the tree representation");
    auto* s = new nbh_sim();
    try {
        s->cfg = *config;
        s->n = count;
        s->key_bytes = config->bits_per_axis == 10 ? 4 : 8;
        s->passes = config->bits_per_axis == 10 ? Layout<10>::sort_passes
                                                : Layout<21>::sort_passes;
        if (hip_stream) {
            s->stream = static_cast<hipStream_t>(hip_stream);
        } else {
            NBH_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
            s->own_stream = true;
        }
        alloc_state(s);
        for (auto& a : s->acc)
            for (auto& c : a)
                launch_fill_f64(s->stream, c.p, 0.0, count);
    } catch (...) {
        delete s;
        throw;
    }
    return s;
}

void validate_config(const nbh_sim_config* c)
{
    if (!c)
        fail(NBH_INVALID_ARGUMENT, "config must not be null");
    if (!std::isfinite(c->time_step) || c->time_step <= 0.0)
        fail(NBH_INVALID_ARGUMENT, "Simulation timeStep must be finite and positive");
    if (!std::isfinite(c->theta) || c->theta < 0.0)
        fail(NBH_INVALID_ARGUMENT, "Simulation theta must be finite and non-negative");
    if (!std::isfinite(c->softening) || c->softening < 0.0)
        fail(NBH_INVALID_ARGUMENT,
             "Simulation softening must be finite and non-negative");
    if (!std::isfinite(c->gravitational_constant) || c->gravitational_constant <= 0.0)
        fail(NBH_INVALID_ARGUMENT, "Gravitational constant must be finite and positive");
    if (c->precision != NBH_PRECISION_FP64 && c->precision != NBH_PRECISION_MIXED)
        fail(NBH_INVALID_ARGUMENT, "precision must be NBH_PRECISION_FP64 or _MIXED");
    if (c->bits_per_axis != 10 && c->bits_per_axis != 21)
        fail(NBH_INVALID_ARGUMENT, "bits_per_axis must be 10 or 21");
}

} // namespace nbh

using namespace nbh;

extern "C" {

int nbh_phase_count(void) { return PH_COUNT; }

const char* nbh_phase_name(int phase)
{
    return (phase >= 0 && phase < PH_COUNT) ? kPhaseNames[phase] : "";
}

int nbh_sim_create(nbh_sim** out, const nbh_sim_config* config, size_t count,
                   const double* h_mass, const double* h_x, const double* h_y,
                   const double* h_z, const double* h_vx, const double* h_vy,
                   const double* h_vz, void* hip_stream)
{
    return guard([&] {
        if (!out)
            fail(NBH_INVALID_ARGUMENT, "output pointer must not be null");
        validate_config(config);
        if (count != 0 && (!h_mass || !h_x || !h_y || !h_z || !h_vx || !h_vy || !h_vz))
            fail(NBH_INVALID_ARGUMENT, "particle arrays must not be null");
        nbh_sim* s = sim_new(config, count, hip_stream);
        try {
            const size_t bytes = count * sizeof(double);
            const double* src[7] = {h_mass, h_x, h_y, h_z, h_vx, h_vy, h_vz};
            nbh::DevBuf<double>* dst[7] = {&s->m, &s->x, &s->y, &s->z,
                                           &s->vx, &s->vy, &s->vz};
            for (int k = 0; k < 7; ++k)
                NBH_HIP(hipMemcpyAsync(dst[k]->p, src[k], bytes,
                                       hipMemcpyHostToDevice, s->stream));
            NBH_HIP(hipStreamSynchronize(s->stream));
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

int nbh_sim_create_device(nbh_sim** out, const nbh_sim_config* config,
                          size_t count, const double* d_mass, const double* d_x,
                          const double* d_y, const double* d_z,
                          const double* d_vx, const double* d_vy,
                          const double* d_vz, void* hip_stream)
{
    return guard([&] {
        if (!out)
            fail(NBH_INVALID_ARGUMENT, "output pointer must not be null");
        validate_config(config);
        if (count != 0 && (!d_mass || !d_x || !d_y || !d_z || !d_vx || !d_vy || !d_vz))
            fail(NBH_INVALID_ARGUMENT, "particle arrays must not be null");
        nbh_sim* s = sim_new(config, count, hip_stream);
        try {
            const size_t bytes = count * sizeof(double);
            const double* src[7] = {d_mass, d_x, d_y, d_z, d_vx, d_vy, d_vz};
            nbh::DevBuf<double>* dst[7] = {&s->m, &s->x, &s->y, &s->z,
                                           &s->vx, &s->vy, &s->vz};
            for (int k = 0; k < 7; ++k)
                NBH_HIP(hipMemcpyAsync(dst[k]->p, src[k], bytes,
                                       hipMemcpyDeviceToDevice, s->stream));
            NBH_HIP(hipStreamSynchronize(s->stream));
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

int nbh_sim_destroy(nbh_sim* sim)
{
    return guard([&] { delete sim; });
}

int nbh_sim_initialize(nbh_sim* s)
{
    return guard([&] {
        if (!s)
            fail(NBH_INVALID_ARGUMENT, "simulation must not be null");
        if (s->state != 0)
            fail(NBH_LOGIC_ERROR, "Simulation can only be initialized once");
        sim_initialize(s);
    });
}

int nbh_sim_step(nbh_sim* s)
{
    return guard([&] {
        if (!s)
            fail(NBH_INVALID_ARGUMENT, "simulation must not be null");
        if (s->state != 1)
            fail(NBH_LOGIC_ERROR, "Simulation must be initialized before stepping");
        sim_step(s);
    });
}

int nbh_sim_run(nbh_sim* s, size_t steps)
{
    if (!s) {
        set_last_error(0, "simulation must not be null");
        return NBH_INVALID_ARGUMENT;
    }
    if (s->state != 1) {
        set_last_error(0, "Simulation must be initialized before running");
        return NBH_LOGIC_ERROR;
    }
    for (size_t k = 0; k < steps; ++k) {
        const int r = nbh_sim_step(s);
        if (r != NBH_OK)
            return r;
    }
    return NBH_OK;
}

double nbh_sim_time(const nbh_sim* s) { return s ? s->time : 0.0; }
size_t nbh_sim_step_number(const nbh_sim* s) { return s ? s->steps : 0; }

int nbh_sim_particles(nbh_sim* s, nbh_particles_view* out)
{
    return guard([&] {
        if (!s || !out)
            fail(NBH_INVALID_ARGUMENT, "arguments must not be null");
        out->count = s->n;
        out->mass = s->m.p;
        out->x = s->x.p;
        out->y = s->y.p;
        out->z = s->z.p;
        out->vx = s->vx.p;
        out->vy = s->vy.p;
        out->vz = s->vz.p;
    });
}

int nbh_sim_accelerations(nbh_sim* s, double** d_ax, double** d_ay,
                          double** d_az)
{
    return guard([&] {
        if (!s || !d_ax || !d_ay || !d_az)
            fail(NBH_INVALID_ARGUMENT, "arguments must not be null");
        *d_ax = s->acc[s->cur][0].p;
        *d_ay = s->acc[s->cur][1].p;
        *d_az = s->acc[s->cur][2].p;
    });
}

int nbh_sim_structure_view(nbh_sim* s, nbh_sim_structure* out)
{
    return guard([&] {
        if (!s || !out)
            fail(NBH_INVALID_ARGUMENT, "arguments must not be null");
        if (s->state == 0)
            fail(NBH_LOGIC_ERROR, "Simulation has no structure before initialize");
        sim_fill_views(s);
        NBH_HIP(hipStreamSynchronize(s->stream));
        std::memset(out, 0, sizeof(*out));
        out->n_particles = static_cast<int>(s->n);
        out->n_groups = s->ngroups;
        out->n_octree_nodes = s->nnodes;
        out->bits_per_axis = s->cfg.bits_per_axis;
        out->sorted_keys = s->keys.p;
        out->sorted_indices = s->idx.p;
        out->unique_keys = s->unique.p;
        out->group_first = s->gfirst.p;
        out->group_count = s->gcount.p;
        out->tree.n_leaves = s->ngroups;
        out->tree.n_internal_nodes = s->ngroups - 1;
        out->tree.left = s->left.p;
        out->tree.right = s->right.p;
        out->tree.parent = s->tparent.p;
        out->tree.range_first = s->rfirst.p;
        out->tree.range_last = s->rlast.p;
        out->tree.prefix_length = s->plen.p;
        out->octree = s->octree_view();
        out->radix_level = s->plevel.p;
        out->edge_node_count = s->pcount.p;
        out->edge_node_offset = s->poffset.p;
        out->target_order = s->horder.p;
        out->box = s->box;
    });
}

int nbh_sim_set_timing(nbh_sim* s, int enabled)
{
    return guard([&] {
        if (!s)
            fail(NBH_INVALID_ARGUMENT, "simulation must not be null");
        s->timing = enabled != 0;
    });
}

int nbh_sim_phase_times(nbh_sim* s, double* ms, int capacity)
{
    if (!s || !ms)
        return NBH_INVALID_ARGUMENT;
    const int k = capacity < PH_COUNT ? capacity : PH_COUNT;
    for (int p = 0; p < k; ++p)
        ms[p] = s->phase_ms[p];
    return PH_COUNT;
}

int nbh_sim_count_interactions(nbh_sim* s, uint64_t* pp, uint64_t* pn,
                               uint64_t* wave_visits)
{
    return guard([&] {
        if (!s || !pp || !pn)
            fail(NBH_INVALID_ARGUMENT, "arguments must not be null");
        if (s->state != 1)
            fail(NBH_LOGIC_ERROR, "Simulation must be initialized");
        launch_fill_u64(s->stream, s->counters.p, 0ull, 4);
        forces(s, s->cur, s->counters.p);
        unsigned long long h[4];
        NBH_HIP(hipMemcpyAsync(h, s->counters.p, sizeof(h),
                               hipMemcpyDeviceToHost, s->stream));
        NBH_HIP(hipStreamSynchronize(s->stream));
        *pp = h[0];
        *pn = h[1];
        if (wave_visits)
            *wave_visits = h[2];
    });
}

} // extern "C"
