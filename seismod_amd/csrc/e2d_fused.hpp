// e2d_fused.hpp -- one-pass FP32 time step of the 2D elastic split-PML
// solver, for one shot or a batch of independent shots.
//
// One launch advances every listed shot by one time step: the velocity
// phase (update_vx/update_vz and pml_v) and the stress phase
// (update_txxzz/update_txz and pml_t), then add_src and receiver sampling,
// as e_iso.c:589-669 orders them.
//
// Work unit: one wavefront owns a 62-cell z segment of a chunk of x rows
// (the chunk length is a per-shot launch parameter).  Lanes 1..62 own cells J0..J0+61; lanes 0 and 63 hold the z
// neighbours.  The wave marches in x; each row of old stress and velocity
// is read once, the velocity of the halo row and halo lanes is recomputed
// from the old state, and z neighbours move between lanes by shuffles.  The
// state is double-buffered (step parity), so no wave reads a value another
// wave has already advanced.  The PML stress split fields are read and
// written only by their owning lane and stay single-buffered.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <tuple>
#include <vector>

struct sm_dbuf;

namespace sm {

constexpr int E2D_OWN = 62;        // cells a wave owns along z
constexpr int E2D_WAVES = 4;       // independent waves per 256-thread block

// State arrays of one parity buffer (each NX*NZ floats).
enum { E2_VX, E2_VZ, E2_TXX, E2_TZZ, E2_TXZ, E2_VXX, E2_VXZ, E2_VZX, E2_VZZ, E2_NSTATE };
// Stress split arrays (single buffer).
enum { E2_TXXX, E2_TXXZ, E2_TZZX, E2_TZZZ, E2_TXZX, E2_TXZZ, E2_NSPLIT };
// Material arrays, dt/h folded in: dt*b/h, dt*bm/h, dt*la/h, dt*l2m/h, dt*mu/h.
enum { E2_CB, E2_CBM, E2_CLA, E2_CL2M, E2_CMU, E2_NMAT };

// Descriptor of one shot, read by the kernel from device memory.
struct E2dShot {
    int NX, NZ;          // padded grid
    int np, nx, nz;      // PML width and interior extent
    int ntz, nch, chunk; // z tiles, x chunks, rows per chunk
    int item0;           // first work item of this shot in the launch
    int nsrc, nrec;
    int it0;             // shot step when the run started
    int nsteps;          // steps of this run
    int rcv0, rmax;      // trace row = it - rcv0, recorded if in [0, rmax)
    unsigned N;          // NX*NZ
    float *st[2];        // E2_NSTATE arrays per parity
    float *ss;           // E2_NSPLIT arrays
    const float *mat;    // E2_NMAT arrays
    const float *dmp;    // (A,B) pairs: x-int[NX], x-half[NX], z-int[NZ], z-half[NZ]
    const int4 *src;     // (padded cell, type, length, fct offset), sorted by cell
    const int *srow;     // [NX+1]: first source at or after row I
    const float *sfct;   // dt*fct
    const int4 *rcv;     // (padded cell, comp, trace column, 0), sorted by cell
    const int *rrow;     // [NX+1]: first receiver at or after row I
    float *trace;        // [rmax][nrec]
};

// Work items (waves) of a shot for a rows-per-chunk value, rounded up to
// whole blocks so that no block spans two shots.
inline int e2d_items(int NX, int NZ, int chunk, int *ntz, int *nch)
{
    *ntz = (NZ + E2D_OWN - 1) / E2D_OWN;
    *nch = (NX + chunk - 1) / chunk;
    const int n = *ntz * *nch;
    return (n + E2D_WAVES - 1) / E2D_WAVES * E2D_WAVES;
}

// Runs a list of shots for their own step counts on one stream: one launch
// per step over the shots still running, long stretches replayed as
// captured hipGraphs.  The caller fills every field of each shot except
// item0.
class E2dRunner {
public:
    ~E2dRunner();
    int run(std::vector<E2dShot> shots, hipStream_t st, int graph_steps);
    // the same without allocating: the table goes to the device's static
    // scratch array (see e2d_scratch_init), steps are plain launches
    int run_static(std::vector<E2dShot> shots, int device, hipStream_t st);
    size_t device_bytes() const;
    // with timing on, each run is bracketed by HIP events on its stream;
    // totals accumulate (graph replays count each of their step launches)
    bool timing = false;
    double ms_total = 0.0;
    size_t launches = 0;
    // kernel form: rows of the next iteration loaded per lane (false) or
    // staged through LDS in 16-byte chunks (true); results are identical
    bool staged = false;

private:
    void launch_step(int nb, hipStream_t st, const E2dShot *tab, const int *bshot, int nitems,
                     const int *base, int k);
    int ensure(size_t bytes);
    int graph_for(int nact, int nitems, int g, hipStream_t st, hipGraphExec_t *out);
    static int plan(std::vector<E2dShot> &shots, std::vector<int> &pref,
                    std::vector<unsigned char> &host, size_t &tab_bytes);
    int launch_all(const std::vector<E2dShot> &shots, const std::vector<int> &pref,
                   const void *table, size_t tab_bytes, hipStream_t st, int graph_steps,
                   size_t *count);
    void drop_graphs();
    void *table_ = nullptr;     // shot descriptors, then the block -> shot map
    size_t table_cap_ = 0;      // bytes
    size_t tab_bytes_ = 0, last_tab_bytes_ = 0;
    bool last_staged_ = false;
    int *base_ = nullptr;
    // instantiated graphs by (active shots, work items, steps per graph),
    // with the tick of their last use; at most MAX_GRAPHS live, the least
    // recently used one destroyed (after its stream drains) to make room
    struct Cached {
        hipGraphExec_t exec;
        uint64_t used;
    };
    std::map<std::tuple<int, int, int>, Cached> graphs_;
    uint64_t tick_ = 0;

public:
    static constexpr size_t MAX_GRAPHS = 16;
    size_t graphs_live() const { return graphs_.size(); }
    size_t graphs_built = 0;    // instantiations since creation
};

// Address of the allocation-free batch scratch on `device` (which must be
// current), looked up once per device; called by sm_ctx_create.
int e2d_scratch_init(int device);
size_t e2d_scratch_bytes();

}  // namespace sm
