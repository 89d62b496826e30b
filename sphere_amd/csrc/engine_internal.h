// engine_internal.h -- the handle behind the C-ABI, shared by the step
// engine (engine.hip) and the adjoint (adjoint.hip).
#ifndef SPHERE_ENGINE_INTERNAL_H_
#define SPHERE_ENGINE_INTERNAL_H_

#include <hip/hip_runtime.h>

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "binio.h"
#include "dem_internal.h"

namespace sph {
struct AdjointState;  // adjoint.hip
}

struct sph_dem {
    sph::HostState hs;
    sph::DemDev d{};
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    // Captured steps, [lead][parity]: parity = the set holding the private
    // records at the step's start; lead = 1 runs the previous step's wall
    // update on a side stream beside the cell list and contact search.
    hipGraph_t graph[2][2] = {};
    hipGraphExec_t gexec[2][2] = {};
    int own_cur = 0;  // 0: the current private records are in b, 1: in a
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    bool use_graph = true;
    int precision = SPH_PREC_FP64;
    int profiling = 0;
    int verbosity = 0;
    sph_timing_t timing{};
    hipEvent_t ev[8] = {};
    bool events_made = false;
    std::string err;
    uint64_t steps_done = 0;
    int device = 0;
    // automatic contact-search choice (search mode 0)
    bool search_auto = true;
    bool force_auto = true;  // k_force variant chosen by timing (sph_set_force_kernel 0)
    uint64_t steps_since_tune = 0;
    bool tuned = false;
    // checkpoints, tape and adjoint buffers of sph_adjoint_* (adjoint.hip)
    std::unique_ptr<sph::AdjointState, void (*)(sph::AdjointState*)> adj{nullptr, nullptr};
    // ensemble (sph_ens_*): the members' host states (scalars, walls and,
    // until upload, particles); hs then holds their concatenation and
    // sys_host the table behind d.systems
    std::vector<sph::HostState> members;
    std::vector<sph::Sys> sys_host;
    bool blocks_auto = false;  // tune threads per block with the search kernel
    int check_vals = 1;        // as sph_open was called (sph_readbin applies it again)
};

namespace sph {

// Work inserted into one step: before_force runs after the contact search
// (the step's sorted state and contact slots are final, the springs are not
// yet updated), after_force after the force kernel.
struct StepHooks {
    std::function<void(const DemDev&, hipStream_t)> before_force, after_force;
};

// Enqueues one step whose private records are read from set `parity` (0: b,
// 1: a) and written to the other; `lead` runs the previous step's wall
// update on the side stream first, `tail` ends with this step's.
void enqueue_step(sph_dem* h, bool timed, int parity, bool lead, bool tail, const StepHooks* hooks = nullptr);

// Runs a few ordinary steps with each contact-search kernel and keeps the
// fastest (search mode 0); with blocks_auto then also the threads per block
// of the search and force kernels.  *ran receives the number of steps run.
int tune_search(sph_dem* h, uint64_t* ran = nullptr);

// Device copy of every array a step reads or a caller can observe (both
// buffer sets, contact history, cell keys with their counts and ranks, the
// last step's outputs, walls, bonds, time and error flags), with the host
// counters that go with it.  Buffers are allocated by the first save_state
// and freed with the checkpoint.
struct Checkpoint {
    std::vector<void*> bufs;
    int own_cur = 0;
    double t_cur = 0.0;
    uint64_t steps_done = 0;
    Checkpoint() = default;
    Checkpoint(const Checkpoint&) = delete;
    Checkpoint& operator=(const Checkpoint&) = delete;
    Checkpoint(Checkpoint&& o) noexcept
        : bufs(std::move(o.bufs)), own_cur(o.own_cur), t_cur(o.t_cur), steps_done(o.steps_done) {
        o.bufs.clear();
    }
    ~Checkpoint();
};
int save_state(sph_dem* h, Checkpoint& c);
// Copies the checkpoint back and requests a neighbour-list rebuild.
int restore_state(sph_dem* h, const Checkpoint& c);

// When the kernel choice is automatic and due (never tuned, or 4096 steps
// since the last tuning): saves the state, runs tune_search on it and
// restores it, so the caller's steps all run with the chosen kernels and
// tuning leaves no trace in the results.
int autotune(sph_dem* h);

}  // namespace sph

#endif
