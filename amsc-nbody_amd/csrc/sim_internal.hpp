// sim_internal.hpp -- state of the Simulation object, shared by the step,
// energy and adjoint code.
#pragma once

#include "nbh_internal.hpp"

#include <cstddef>

namespace nbh {

const char* traversal_message(int code);
const char* stage_message(int word, int code);

// Start of each level's internal nodes in the level-grouped node list
// (s[L + 1] ends level L; s[kMaxLevels] is the internal node total).
struct LevelStarts {
    int s[kMaxLevels + 1];
};
// The level starts from the plan scan's per-level counts (one counter per
// kLevelCountStride ints, the root excluded).
LevelStarts level_starts(const int* level_count);
// The internal nodes grouped by level: order[start[L] .. start[L + 1]) are
// the octree nodes of level L (the root first), from the plan.  cursor:
// kMaxLevels * kLevelCountStride zeroed ints; n_leaves the host's group
// count.
void launch_level_scatter(hipStream_t s, const int* d_nleaves, int n_leaves,
                          const int* plevel, const int* pcount, const int* poffset,
                          const int* level_count, int* cursor, int* order);
// Mass / COM, subtree sizes and the walk layout over the level-grouped
// internal nodes and the leaves (one per Morton group).
template <int BITS>
void launch_level_passes(hipStream_t s, const nbh_octree_view& o, const LevelStarts& st,
                         bool mixed, const double4* rec64, const int4* rec32, double theta,
                         const double* box, bool need_center, double eps,
                         const LevelArrays& la, const int* order, const int* poffset,
                         const int* pcount, int n_groups, double G, Walk64Rec* w64,
                         double4* w64center, int4* w32node, int4* w32center, int* wcount,
                         hipEvent_t between, uint32_t* horder, const uint8_t* htab);
// The leaves' {pre, curve offset} words of the last layout (the layout
// passes write them for the internal nodes only).
void launch_leaf_positions(hipStream_t s, const nbh_octree_view& o, const int* poffset,
                           const int* pcount, int n_groups, const LevelArrays& la);
void launch_walk_sentinel(hipStream_t s, int4* w32node, Walk64Rec* w64, int cap);
template <int BITS>
void launch_octree_views(hipStream_t s, const nbh_octree_view& o, const LevelArrays& la,
                         const double4* rec64, const double* box);

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



// The mid-step readback, one copy: the device keeps the same layout at the
// end of its per-step cleared words (errors, level counters and scatter
// cursors are cleared; the counts and the box are written every step), the
// host a pinned copy.
constexpr int kErrorWords = 64;
struct Readback {
    int errors[kErrorWords]; // ERR_WORD_COUNT used
    // internal octree nodes per level (root excluded), one line per level
    int level_count[kMaxLevels * kLevelCountStride];
    int level_cursor[kMaxLevels * kLevelCountStride]; // level scatter cursors
    int ngroups;
    int nnodes;
    int pad[2];
    double box[5]; // centre x, y, z, side, largest |m|
};
constexpr size_t kReadbackClearedWords =
    offsetof(Readback, ngroups) / sizeof(int);
static_assert(ERR_WORD_COUNT <= kErrorWords, "error words fit the readback");
static_assert(offsetof(Readback, box) % 8 == 0, "box is 8-byte aligned");

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
    nbh::DevBuf<int> pplen; // per radix node: its parent's prefix length
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
    // the device side of the mid-step readback (inside `zero`): errors,
    // group and node counts, level starts, box
    nbh::Readback* mid = nullptr;
    double* box = nullptr;        // mid->box: centre x, y, z, side, largest |m|
    // small device scalars
    nbh::DevBuf<double> partials; // bbox partials
    nbh::DevBuf<int> scalars;     // [3] wcount ([0..2] unused)
    nbh::DevBuf<unsigned long long> counters;
    nbh::DevBuf<int> walkq;       // mixed traversal work-queue counters
    // traversal schedule (mixed walks, and the FP64 force VJP of few
    // packets): the queue order of the units, then each unit's cost in the
    // last walk (zeroed at creation)
    nbh::DevBuf<uint32_t> psched;
    // mixed traversal split into subtree units (few packets only): the
    // units' position ranges and their partial sums
    nbh::DevBuf<int> wsplit;
    nbh::DevBuf<double> wpart;

    // octree (capacity grows)
    size_t mcap = 0;
    int nnodes = 0;
    int ngroups = 0;
    nbh::DevBuf<int> children, oparent, level, ofirst, ocount, pending,
        child_count, order;
    nbh::DevBuf<unsigned char> prefix;
    nbh::DevBuf<double> mass, comx, comy, comz, cenx, ceny, cenz, half;
    nbh::DevBuf<nbh::Walk64Rec> w64; // FP64 walk records
    nbh::DevBuf<double4> w64center;
    // per-node state of the level passes
    nbh::DevBuf<double4> sum4;
    nbh::DevBuf<int4> aux;
    nbh::DevBuf<double> crit;
    nbh::DevBuf<int2> link, pos;
    nbh::LevelStarts starts{};       // level starts of the last rebuild
    bool views_fresh = false;        // reference-layout node arrays match sum4
    nbh::DevBuf<int4> w32node, w32center;
    // the traversal's target order along the Hilbert curve
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

    nbh::LevelArrays level_arrays()
    {
        return nbh::LevelArrays{sum4.p, aux.p, crit.p, link.p, pos.p};
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
