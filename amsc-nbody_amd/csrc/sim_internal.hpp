// sim_internal.hpp -- state of the Simulation object, shared by the step,
// energy and adjoint code.
#pragma once

#include "nbh_internal.hpp"

namespace nbh {

const char* traversal_message(int code);
const char* stage_message(int word, int code);

template <int BITS>
void launch_level_passes(hipStream_t s, const nbh_octree_view& o, int max_level,
                         bool mixed, const double4* rec64, const int4* rec32,
                         double theta, const double* box,
                         bool need_center, double eps, int* size, int* local,
                         double* crit, int* pre, int* order, int* level_info,
                         double4* sum4, const int* poffset, const int* pcount,
                         int n_groups, double G, Walk64Rec* w64, double4* w64center,
                         int4* w32node,
                         int4* w32link, int4* w32center, int* wcount,
                         hipEvent_t between, int* hcount = nullptr,
                         int* hoffset = nullptr, uint32_t* horder = nullptr,
                         const uint8_t* htab = nullptr);
// ints of the level-pass bookkeeping array (level starts + per-block counts)
size_t level_info_ints();
void launch_walk_sentinel(hipStream_t s, int4* w32node, Walk64Rec* w64, int cap);
template <int BITS>
void launch_octree_views(hipStream_t s, const nbh_octree_view& o,
                         const double4* sum4, const double* box);

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { reset(); }
    void reset()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count)
    {
        reset();
        if (count == 0)
            count = 1;
        NBH_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
        n = count;
    }
    void ensure(size_t count)
    {
        if (count > n)
            alloc(count + count / 4);
    }
};

enum Phase : int {
    PH_KICK_DRIFT_BBOX = 0,
    PH_ENCODE,
    PH_SORT,
    PH_GATHER,
    PH_GROUPS,
    PH_KARRAS,
    PH_PLAN,
    PH_OCTREE,
    PH_COM,
    PH_WALK,
    PH_TRAVERSAL,
    PH_KICK,
    PH_COUNT
};



// Host copy of the mid-step readback (pinned).
struct Readback {
    int ngroups;
    int nnodes;
    int errors[ERR_WORD_COUNT];
    int pad;
    double box[5]; // centre x, y, z, side, largest |m|
};

} // namespace nbh

struct nbh_sim {
    nbh_sim_config cfg{};
    size_t n = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int state = 0; // 0 created, 1 ready, 2 failed
    double time = 0.0;
    size_t steps = 0;
    bool timing = false;
    int key_bytes = 8;
    int passes = 8;

    // particle state, original order
    nbh::DevBuf<double> m, x, y, z, vx, vy, vz;
    nbh::DevBuf<double> acc[2][3];
    int cur = 0;

    // build buffers sized for n
    nbh::DevBuf<unsigned char> keys, keys_alt, unique;
    nbh::DevBuf<uint32_t> idx, idx_alt;
    nbh::DevBuf<double4> rec64;
    nbh::DevBuf<double4> pos4; // packed {x, y, z, m}, original order
    nbh::DevBuf<int4> rec32;
    nbh::DevBuf<int> gfirst, gcount;
    nbh::DevBuf<int> left, right, tparent, rfirst, rlast, plen;
    nbh::DevBuf<int> plevel, pcount, poffset;
    // per-step cleared words: hist | sort status | group scan | plan scan |
    // error words
    nbh::DevBuf<uint32_t> zero;
    size_t zero_words = 0;
    uint32_t* hist = nullptr;
    uint32_t* sort_status = nullptr;
    unsigned long long* group_status = nullptr;
    unsigned long long* plan_status = nullptr;
    int* errors = nullptr;
    // small device scalars
    nbh::DevBuf<double> box;      // centre x, y, z, side, largest |m|
    nbh::DevBuf<double> partials; // bbox partials
    nbh::DevBuf<int> scalars;     // [1] ngroups [2] nnodes [3] wcount ([0] unused)
    nbh::DevBuf<unsigned long long> counters;
    nbh::DevBuf<int> walkq;       // mixed traversal work-queue counters
    // mixed traversal split into subtree units (few packets only): the
    // units' position ranges and their partial sums
    nbh::DevBuf<int> wsplit;
    nbh::DevBuf<double> wpart;

    // octree (capacity grows)
    size_t mcap = 0;
    int nnodes = 0;
    int ngroups = 0;
    nbh::DevBuf<int> children, oparent, level, ofirst, ocount, pending,
        child_count, size, local, pre, order, level_info;
    nbh::DevBuf<unsigned char> prefix;
    nbh::DevBuf<double> mass, comx, comy, comz, cenx, ceny, cenz, half, crit;
    nbh::DevBuf<nbh::Walk64Rec> w64; // FP64 walk records
    nbh::DevBuf<double4> w64center;
    nbh::DevBuf<double4> sum4; // packed {com, mass} per node (level passes)
    bool views_fresh = false;  // reference-layout node arrays match sum4
    nbh::DevBuf<int4> w32node, w32center;
    // the traversal's target order along the Hilbert curve
    nbh::DevBuf<int> hcount, hoffset;    // per node
    nbh::DevBuf<uint32_t> horder;        // per particle
    nbh::DevBuf<uint8_t> htab;           // the curve's state tables

    nbh::Readback* rb = nullptr;
    hipEvent_t ev[nbh::PH_COUNT + 1] = {};
    double phase_ms[nbh::PH_COUNT] = {};

    ~nbh_sim()
    {
        if (rb)
            (void)hipHostFree(rb);
        for (auto& e : ev)
            if (e)
                (void)hipEventDestroy(e);
        if (own_stream && stream)
            (void)hipStreamDestroy(stream);
    }

    nbh_octree_view octree_view()
    {
        nbh_octree_view o{};
        o.n_particles = static_cast<int>(n);
        o.n_leaves = ngroups;
        o.n_nodes = nnodes;
        o.root = 0;
        o.children = children.p;
        o.parent = oparent.p;
        o.level = level.p;
        o.prefix = prefix.p;
        o.first_particle = ofirst.p;
        o.particle_count = ocount.p;
        o.mass = mass.p;
        o.com_x = comx.p;
        o.com_y = comy.p;
        o.com_z = comz.p;
        o.center_x = cenx.p;
        o.center_y = ceny.p;
        o.center_z = cenz.p;
        o.half_size = half.p;
        o.pending_children = pending.p;
        o.child_count = child_count.p;
        return o;
    }

    void mark(int phase)
    {
        if (timing)
            NBH_HIP(hipEventRecord(ev[phase], stream));
    }
};


namespace nbh {

bool mixed(const nbh_sim* s);
void validate_config(const nbh_sim_config* c);
void alloc_state(nbh_sim* s);
// Rebuild at the current positions (after a fused kick-drift when
// kick_drift): Morton keys through the walk layout.  Synchronises once.
void do_rebuild(nbh_sim* s, bool kick_drift);
// Forces of the current tree into acceleration buffer `buffer`.
void forces(nbh_sim* s, int buffer, unsigned long long* counters);
// Ends a step: reads the device error words (synchronises) and the phase
// timings.
void finish(nbh_sim* s);
// Fills the reference-layout octree arrays (mass, centre of mass, cell
// geometry, child counts) of the last rebuild, once per rebuild.
void sim_fill_views(nbh_sim* s);
// Allocates a simulation of `count` particles with zero accelerations and
// unset particle state (validates the configuration and the count).
nbh_sim* sim_new(const nbh_sim_config* config, size_t count, void* hip_stream);
// Simulation::initialize / Simulation::step without the state checks; a
// failure marks the simulation failed and rethrows.
void sim_initialize(nbh_sim* s);
void sim_step(nbh_sim* s);

} // namespace nbh
