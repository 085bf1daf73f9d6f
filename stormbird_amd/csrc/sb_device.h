// Device-side data layout of one lifting-line instance and the kernel entry
// points that operate on it (implemented in sb_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "sb_common.h"
#include "sb_mlp.h"
#include "sb_wind.h"

namespace sb {

// Packed per-step geometry (uploaded from the host after the rigid-body update).
// Offsets are in doubles from the start of the geometry buffer.
struct GeoLayout {
    int ctrl;          // [N][3] ctrl points (global)
    int rel_g;         // [N][3] span line vectors (global)
    int dir_g;         // [N][3] normalized span line vectors (global)
    int chord_g;       // [N][3] chord vectors (global)
    int chord_len;     // [N]
    int local_len;     // [N] length of the local span line
    int rel_local;     // [N][3] span line vectors (local)
    int rel_out;       // [N][3] span line vectors in the output coordinate system
    int dir_out;       // [N][3] their directions
    int ctrl_out;      // [N][3] ctrl points in the output coordinate system
    int chord_out;     // [N][3] chord vectors in the output coordinate system
    int is_virtual;    // [N] 1.0 for virtual strips
    int span_pts;      // [P][3] span points (global)
    int end_g;         // [N][3] span line end points (global)
    int chord_span;    // [P][3] chord vectors at span points
    int rbm;           // [18]: translation, rotation, v_lin, v_ang, (unused 6)
    int body_axes;     // [9]: rotated unit x, y, z axes (body-fixed output)
    int total;
};

// Instance descriptor passed by value to every kernel.
// Doubles per panel of the near-field record (Inst::segs): 12 corner coordinates
// and 4 x (unit direction, length) of the ring edges.
constexpr int NEARD = 28;

struct Inst {
    int N, W, P, M;     // strips, wings, span points, wake rows (panels per strip)
    int nz_rows;        // leading panel rows that may hold nonzero strength
    int symmetry;       // 0 none, 1 X, 2 Y, 3 Z
    int neglect_self;   // skip same-wing pairs in the frozen wake
    int first_dir;      // 0 chord, 1 freestream, 2 actual velocity
    int steps_completed;
    int output_body;
    int viscous_lift_in_circulation;
    int vc_mode;        // velocity correction: 0 none, 1 max ratio, 2 fixed magnitude
    int max_iterations;
    long end_index;     // wake points [0, end_index) receive induced velocities
    double far_ratio, first_rel, last_rel, damping, rep_chord;
    double solver_damping, residual_tol, strength_tol, vc_ratio, correction_factor;
    double density;
    double dt;           // time step (batched launches read it here)
    // quasi-steady wake and Linearized solver
    int quasi;           // horseshoe frozen wake, no shed wake (M == 0)
    int solver_kind;     // 0 SimpleIterative, 1 Linearized
    int start_linearized;     // SimpleIterative starts from a default Linearized solve
    int lin_disable_viscous;  // Linearized: skip the viscous circulation correction
    int lin_ivc;              // Linearized: 1 recompute the induced velocity after it
    double hs_wake_length;    // horseshoe trailing-leg length
    double hs_core;           // horseshoe viscous core length
    double* lin;         // Linearized scratch when the matrix does not fit in shared memory
                         // (linearized_global_doubles(N) doubles)
    double* lin_gamma;   // [N] Linearized circulation that starts SimpleIterative

    double* pts;         // current wake points [P*(M+1)][3]
    double* pts_next;    // next wake points (double buffer)
    double* vel;         // velocity at wake points [P*(M+1)][3]
    double* str;         // current strengths [N*M]
    double* str_next;    // next strengths (double buffer)
    const double* core;  // viscous core length [N*M]
    const double* corners;  // optional explicit corners [N*M][12] (standalone panel sets)
    PanelRec* rec;       // far-field panel records [N*M]
    double* segs;        // optional [N*M][NEARD]: the ring's four corners (ring order), then per
                         // ring edge e (corner e to e+1) its unit direction and length, the
                         // target-independent part of the viscous core term; null: computed per pair
    const int* col;      // [N] span-point column of strip i's first corner (i + wing index)
    const int* wing_of;  // [N]
    const int* span_src; // [P][2] ctrl-point indices mapped to span point j
    const FoilParams* foil;  // [W]
    const double* geo;   // packed geometry (GeoLayout)
    GeoLayout gl;
    const double* fs;    // freestream [N + P*(M+1)][3]
    double* felt;        // felt ctrl velocity [N][3]
    double* felt_span;   // felt span-point velocity [P][3]
    double* fixed;       // fixed induced velocity at ctrl points [N][3]
    double* F;           // unit-strength factors [N(cols j)][N(rows i)][3]
    double* prev_gamma;  // [N] previous circulation (initial guess)
    double* hist;        // flow-derivative history [2][N][3]
    double* scal;        // scalars: [0..2] mean ctrl freestream, [3] fd count, [4] singular-matrix row + 1
    double* out;         // packed step output (OutLayout)
};

// Packed step output, offsets in doubles.
struct OutLayout {
    SB_HD static long gamma(int) { return 0; }
    SB_HD static long alpha(int N) { return N; }
    SB_HD static long u(int N) { return 2L * N; }
    SB_HD static long felt(int N) { return 5L * N; }
    SB_HD static long ctrl(int N) { return 8L * N; }
    SB_HD static long sect(int N) { return 11L * N; }                 // [6][N][3]
    SB_HD static long integ_f(int N) { return 29L * N; }              // [6][W][3]
    SB_HD static long integ_m(int N, int W) { return 29L * N + 18L * W; }
    SB_HD static long iters(int N, int W) { return 29L * N + 36L * W; }
    SB_HD static long residual(int N, int W) { return 29L * N + 36L * W + 1; }
    SB_HD static long singular(int N, int W) { return 29L * N + 36L * W + 2; }  // Linearized: failing step + 1, or 0
    SB_HD static long total(int N, int W) { return 29L * N + 36L * W + 3; }
};

enum Mode { MODE_EXACT = 0, MODE_FAST = 1, MODE_FP32_KAHAN = 2 };

// Scratch of the load-balanced (McNaughton) pair-sum launch, owned by one
// simulation: partial sums of split target blocks and their publish flags.
struct McBuffers {
    double* partial = nullptr;    // [cap_jobs][256][3]
    long long* flags = nullptr;   // [cap_jobs], initialised to -1
    long cap_jobs = 0;
    long long epoch = 0;          // incremented per launch; flags compare against it
    int slots[3][2][2] = {};  // resident workgroups per (mode, symmetry, near-field staging)
    // Order of the self-induction targets: slot j * 256 + l of job j holds wake point
    // tmap[slot], or -(i + 1) for a padding slot that evaluates wake point i and stores
    // nothing.  Null: flat order (slot = wake point).  Each target's own panel order is
    // the same in either order, so the results are too.
    int* tmap = nullptr;
    long tmap_jobs = 0;
    int tmap_span = 0;  // span points per wave of the map (0: flat order)
};

// Scratch of the two-pass few-target sum (k_terms + k_fold_terms): one chunk
// of per-pair terms [3][chunk][ld] and the running sums [3][n_targets].
struct FoldBuffers {
    double* terms = nullptr;
    long cap = 0;                 // doubles in `terms`
    double* acc = nullptr;        // [3 * cap_targets]
    long cap_targets = 0;
};

// Kernel launchers (sb_kernels.hip). All enqueue on `st` and return hipError_t.
hipError_t launch_prepare(const Inst& d, int first_step, hipStream_t st);
hipError_t launch_init_wake(const Inst& d, double dt, int use_mean_velocity, const double* velocity3,
                            hipStream_t st);
hipError_t launch_wake_step(const Inst& d, double dt, hipStream_t st);
hipError_t launch_panels(const Inst& d, const double* pts, int shift_strengths, hipStream_t st);
// Fixed velocities at the control points (rows >= 1).  Two-pass sum when `fb`
// holds room for at least one 64-panel chunk, else one wave per target.
hipError_t launch_fixed(const Inst& d, int mode, FoldBuffers* fb, hipStream_t st);
hipError_t launch_factors(const Inst& d, hipStream_t st);
// Quasi-steady factors: F[j][i] = unit velocity of strip j's horseshoe vortex at ctrl point i.
hipError_t launch_horseshoe_factors(const Inst& d, hipStream_t st);
// Linearized solve; as_start = 1 writes only the circulation to d.lin_gamma, with
// the default Linearized settings (start of a SimpleIterative solve).  Needs d.lin
// unless linearized_in_lds(d.N).
hipError_t launch_linearized(const Inst& d, int as_start, hipStream_t st);
// The one-workgroup Gaussian elimination (variant 1-3) on a dense row-major system.
hipError_t launch_gauss_dense(const double* A, const double* b, int N, double* x, int* status, int variant, hipStream_t st);
// The blocked factorization and back substitution on lin ([N + 1][N] column-major, then 8N doubles and N
// ints of workspace) with scal[4] cleared.
hipError_t launch_gauss_blocked(double* lin, double* scal, int N, double* x, int* status, hipStream_t st);
bool linearized_in_lds(int N);
long linearized_global_doubles(int N);
hipError_t launch_solve(const Inst& d, hipStream_t st);
size_t solve_shared_bytes(int N, int W);
hipError_t launch_set_row0(const Inst& d, hipStream_t st);
hipError_t launch_velocity_copy(const Inst& d, hipStream_t st);
// dst[0, 3n) = src[0, 3n); src may be pinned host memory (read over the bus)
hipError_t launch_copy3(const double* src, double* dst, long n, hipStream_t st);
// dst[0, count) = src[0, count); src may be pinned host memory
hipError_t launch_copy_doubles(const double* src, double* dst, long count, hipStream_t st);
hipError_t launch_self_induction(const Inst& d, int mode, McBuffers* mc, hipStream_t st);
// Induced velocity at `targets` from panels [0, n_panels): out (+)= sum.
hipError_t launch_targets(const Inst& d, const double* targets, long n_targets, long n_panels, int mode, double* out,
                          int add, McBuffers* mc, hipStream_t st, const int* tmap = nullptr, long map_jobs = 0);
hipError_t launch_forces(const Inst& d, double dt, hipStream_t st);
hipError_t launch_flow_update(const Inst& d, hipStream_t st);
// Induced velocity at arbitrary targets from panels [0, n_panels): out = sum.
hipError_t launch_induced_targets(const Inst& d, const double* targets, long n_targets, long n_panels,
                                  int mode, double* out, McBuffers* mc, hipStream_t st);
hipError_t launch_count_pairs(const Inst& d, const double* targets, long n_targets, long n_panels,
                              double* partial, long* n_blocks, hipStream_t st);
// Wind-environment freestream at the wake points (after the N control points of d.fs).
hipError_t launch_wind_points(const Inst& d, const WindField& w, hipStream_t st);

// Freestream of one batched instance from its wind field: the control points'
// values come from the host (inflow corrections are applied there), the wake
// points' are evaluated where the points are.
struct WindJob {
    WindField w;
    const double* ctrl;  // [n_ctrl][3]
    const double* pts;   // [n_pts][3] wake points
    double* dst;         // [n_ctrl + n_pts][3]
    long n_ctrl, n_pts;
};
// blk_off[n + 1]: 256-point blocks per instance, prefix sums (device)
hipError_t launch_batch_wind(const WindJob* jobs, const long* blk_off, int n, long blocks, hipStream_t st);
hipError_t launch_libm(int fn, const double* x, const double* y, double* out, long n, hipStream_t st);

// ---------------------------------------------------------------------------
// Many instances per launch (sb_batch).  Every instance runs the same kernel
// sequence as the single-instance step; a flat grid is mapped back to
// (instance, block of that instance) through per-instance block offsets, and
// the pair sums run as persistent workgroups pulling jobs from a queue.
// ---------------------------------------------------------------------------
enum BatchKind { BK_POINTS = 0, BK_SPAN = 1, BK_PANELS = 2, BK_FACTORS = 3, BK_VEL = 4, BK_KINDS = 5 };

struct BatchDev {
    const Inst* inst = nullptr;   // device [n], the descriptors of the current stage
    int n = 0;
    const long* off = nullptr;    // device [BK_KINDS][n + 1]: first flat 256-thread block per instance
    long blocks[BK_KINDS] = {};   // total flat blocks per kind
    int max_threads = 64;         // one-workgroup kernels: max over instances of N rounded to 64
    size_t solve_shm = 0;         // bytes of shared memory for the batched solve
    const int2* fixed_jobs = nullptr;   // (instance, 64-target block), longest first
    int n_fixed_jobs = 0;
    const int2* target_jobs = nullptr;  // (instance, 256-target block) of the free-wake self-induction
    int n_target_jobs = 0;
    unsigned* counters = nullptr;       // device [2] job-queue heads, zero between launches
    int fixed_grid = 0, target_grid = 0;  // persistent workgroups
    int fixed_per = 3;            // kb_fixed: panels per compute wave per block (1, 2, 3)
    int targets_tile = 256;       // kb_targets: panel records per LDS tile (64, 128, 256)
    int stage_near = 0;           // kb_targets: 1 when every self-induced instance has ring-edge
                                  // geometry (near-field data staged with the records)
    int sym = 0;                  // 1 when any instance uses a symmetry plane
    // learned warm start (sb_mlp): per-strip features [n_strips][64], scale, prediction
    const long* strip_off = nullptr;  // device [n + 1]: first strip of each instance
    long n_strips = 0;
    float* feat = nullptr;
    float* scale = nullptr;
    float* pred = nullptr;
    int warm_residual = 0;            // the prediction is the change from the previous circulation
    unsigned* fallbacks = nullptr;    // device [2]: instances that kept their default guess (invalid
                                      // prediction); instances started from the prediction (judged mode)
};

hipError_t launch_batch_prepare(const BatchDev& b, int first_step, hipStream_t st);
hipError_t launch_batch_init_wake(const BatchDev& b, hipStream_t st);
hipError_t launch_batch_wake_step(const BatchDev& b, hipStream_t st);
hipError_t launch_batch_panels(const BatchDev& b, int shift, hipStream_t st);
hipError_t launch_batch_fixed(const BatchDev& b, int mode, hipStream_t st);
hipError_t launch_batch_factors(const BatchDev& b, hipStream_t st);
hipError_t launch_batch_solve(const BatchDev& b, hipStream_t st);
hipError_t launch_batch_velocity_copy(const BatchDev& b, hipStream_t st);
hipError_t launch_batch_self_induction(const BatchDev& b, int mode, hipStream_t st);
hipError_t launch_batch_forces(const BatchDev& b, hipStream_t st);
// Features of every strip (kb_features) and, when `apply` is 1, the network's
// initial circulation written into prev_gamma (after the prepare kernel); with
// `apply` 2 only the prediction is made, for launch_batch_warm_judge.
hipError_t launch_batch_warm_start(const BatchDev& b, sb_mlp* mlp, int apply, hipStream_t st);
// Per instance, the better of the default guess and the prediction by the
// solver's first evaluation (after the factors, before the solve).
hipError_t launch_batch_warm_judge(const BatchDev& b, hipStream_t st);
// Quasi-steady batches: horseshoe factors of every instance, and the
// one-workgroup Linearized solve per instance (variant 1: N <= 128, 2: larger;
// 0 when max_n needs the blocked form, which batches do not run).
hipError_t launch_batch_horseshoe_factors(const BatchDev& b, hipStream_t st);
int batch_linearized_variant(int max_n);
// Shared memory of the variant for instances of ns[0..n) strips (0: one does not fit).
size_t batch_linearized_shared(const int* ns, int n, int variant);
hipError_t launch_batch_linearized(const BatchDev& b, int variant, size_t shm, int as_start, hipStream_t st);
// Singular-pivot flags of every instance into flags[n] (device), cleared in the instances.
hipError_t launch_batch_take_singular(const BatchDev& b, double* flags, hipStream_t st);
// Resident workgroups of the persistent pair-sum kernels (0 fixed, 1 self-induction).
// Resident workgroups of kb_fixed (which 0; variant = panels per wave) or
// kb_targets (which 1; variant = tile length).
int batch_resident_groups(int which, int mode, int sym, int variant, int stage_near);

}  // namespace sb
