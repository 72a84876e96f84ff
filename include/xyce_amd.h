/*
 * xyce_amd — MI355X-native per-Newton hot path of a Xyce transient run.
 *
 * C ABI of libxyce_amd.so.  Plain pointers and sizes only; no exceptions cross
 * it; every entry point returns an int status:
 *    0            success
 *   >0            singular pivot: (pivot index + 1)      (linear entry points)
 *   <0            error (see XG_E_* below)
 * One handle owns one HIP stream; calls on a handle must be serialised by the
 * caller (as Xyce's single-threaded DeviceMgr / Linear::Solver calls are).
 *
 * Reference interfaces replaced (paths relative to the Xyce source root):
 *   device side   Device::updateState / loadDAEVectors / loadDAEMatrices /
 *                 isConverged   src/DeviceModelPKG/Core/N_DEV_Device.h:312-531,
 *                 MOSFET level 1 Master  src/DeviceModelPKG/OpenModels/N_DEV_MOSFET1.C:4241,4343,4587
 *   linear side   Linear::Solver::doSolve(reuse_factors, transpose)
 *                 src/LinearAlgebraServicesPKG/N_LAS_Solver.h:95-107 (KLU via
 *                 N_LAS_AmesosSolver.C:216-478)
 */
#ifndef XYCE_AMD_H
#define XYCE_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define XG_OK 0
#define XG_E_ARG (-1)        /* bad argument / shape */
#define XG_E_HIP (-2)        /* HIP runtime error */
#define XG_E_SETUP (-3)      /* circuit setup failed (e.g. effective length <= 0) */
#define XG_E_NAN (-99)       /* NaN in the solution (N_LAS_AmesosSolver.C:415-424) */
#define XG_E_NOGPU (-10)     /* no HIP device */

#define XG_PRED_WIN 256      /* learned predictor: unknowns per window */
#define XG_PRED_FEAT 1600    /* learned predictor: features per window */

/* Device types in netlist order (dev_type[]). */
#define XG_DEV_MOS1 0
#define XG_DEV_RES 1
#define XG_DEV_CAP 2
#define XG_DEV_VSRC 3
#define XG_DEV_ISRC 4

/* Waveform types. DC: [v]; PULSE: [V1 V2 TD TR TF PW PER] (NaN = default);
 * PWL: [t0 v0 t1 v1 ...] — src/DeviceModelPKG/Core/N_DEV_SourceData.C:1168,1770 */
#define XG_WAVE_DC 0
#define XG_WAVE_PULSE 1
#define XG_WAVE_PWL 2

/* Number of MOSFET level-1 .model parameters per card and instance parameters
 * per device (order: see xyce_amd/circuits.py MODEL_PARAMS / INST_PARAMS). */
#define XG_NMP 30
#define XG_NIP 11

/*
 * Circuit description: a netlist as flat arrays.  Node 0 is ground.
 * Replaces the device/topology construction of Topology::addDevice
 * (src/TopoManagerPKG/N_TOP_Topology.C:191) for the supported device types.
 */
typedef struct xg_circuit_desc {
  int n_nodes;
  int n_dev;            const int* dev_type;   const int* dev_idx;   /* netlist order */
  int n_models;         const double* model_par;                    /* n_models * XG_NMP */
  int n_mos;            const int* mos_nodes;  const int* mos_model; const double* mos_inst; /* 4, 1, XG_NIP per device */
  int n_res;            const int* res_nodes;  const double* res_val;
  int n_cap;            const int* cap_nodes;  const double* cap_val;
  int n_vsrc;           const int* vsrc_nodes; const int* vsrc_wave;
  int n_isrc;           const int* isrc_nodes; const int* isrc_wave;
  int n_wave;           const int* wave_type;  const int* wave_off;  const double* wave_data;
  double gmin;          /* DeviceOptions gmin, default 1e-12 (N_DEV_DeviceOptions.C:90) */
  double temp;          /* Kelvin, default 300.15 */
} xg_circuit_desc;

/* Analysis flags of SolverState::setupSolverInfo (N_DEV_SolverState.C:292-430). */
typedef struct xg_solver_state {
  /* init_fix changes nothing here: the reference applies it only to devices
   * with OFF (N_DEV_MOSFET1.C:2822,2864-2867), which the library does not
   * model; voltage limiting is always on (DeviceOptions voltageLimiterFlag). */
  int dcop, tran, init_jct, init_fix;
  int newton_iter;      /* 0 on the first load of a Newton solve */
  int order;            /* integration order 1 or 2 (OneStep) */
  double h;             /* current time step */
  double time;          /* t_{n+1}: time at which sources are evaluated */
  double tstep, tstop;  /* .TRAN step / stop (PULSE defaults) */
  int tranop;           /* the operating point of a .TRAN (tranopFlag, N_DEV_SolverState.C:334;
                         * N_ANP_AnalysisManager.C:1170): the capacitances are stamped into dQdx and
                         * the charge limiter terms as in a transient (N_DEV_MOSFET1.C:4504,4829) */
} xg_solver_state;

typedef struct xg_circuit xg_circuit;
typedef struct xg_lu xg_lu;
typedef struct xg_batch xg_batch;
typedef struct xg_mlp xg_mlp;

/* ---------------- circuit (device side) ---------------- */

/* Builds the circuit on the current HIP device: model/instance setup
 * (MOSFET1 Model::processParams :3939, Instance::processParams :383,
 * updateTemperature :3487), node ordering and CSR pattern
 * (SerialLSUtil::setupNodeGIDs / setupSolnAndStateGIDs / generateRowColData,
 * N_TOP_SerialLSUtil.C:468,504,620) and the stamp gather maps. */
int xg_circuit_create(const xg_circuit_desc* desc, xg_circuit** out);
int xg_circuit_destroy(xg_circuit* c);

/* sizes[0..5] = n (unknowns), nnz, n_state, n_store, n_mos, n_nodes */
int xg_circuit_sizes(const xg_circuit* c, int* sizes);

/* The Jacobian CSR pattern (rowptr n+1, colind nnz), node -> solution index
 * (-1 = ground), per-MOSFET solution indices (6: d g s b d' s'; n = ground),
 * F/Q stamp slots (17 and 14 per MOSFET; -1 = ground dummy, as
 * EpetraMatrix::operator() N_LAS_EpetraMatrix.C:658), state/store offsets. */
int xg_circuit_pattern(const xg_circuit* c, int* rowptr, int* colind, int* node_gid, int* mos_lid,
                       int* mos_fslot, int* mos_qslot, int* mos_stasto);

int xg_set_solver_state(xg_circuit* c, const xg_solver_state* ss);

/* a2: DeviceMgr::updateState -> MOSFET1 Master::updateState (N_DEV_MOSFET1.C:4241).
 * Host arrays in Xyce layout: x[n], state/store vectors as laid out by
 * xg_circuit_pattern's mos_stasto (NULL only for a circuit without MOSFETs).
 * Writes the MOSFETs' entries of sta_next / sto_next and leaves every other
 * entry as the caller had it (a master writes only its own devices').
 * *limited = !allDevicesConverged (limited may be NULL). */
int xg_update_state(xg_circuit* c, const double* x, const double* sta_cur, const double* sto_cur,
                    double* sta_next, double* sto_next, int* limited);

/* a5+a7: DeviceMgr::loadDAEVectors (N_DEV_MOSFET1.C:4343) plus the linear
 * devices' separateLoad contribution (N_LOA_CktLoader.C:700-787).  Uses the
 * device state of the last xg_update_state.  Each output has n entries and
 * receives this handle's contribution ADDED to what it holds, as the
 * reference's masters "+=" into vectors the loader has zeroed
 * (N_DEV_MOSFET1.C:4393-4411, N_LOA_NonlinearEquationLoader.C:385-390):
 * out = out + contribution, one rounding per entry.  NULL outputs are skipped. */
int xg_load_vectors(xg_circuit* c, double* F, double* Q, double* B, double* dFdxdVp, double* dQdxdVp);

/* a6+a7: DeviceMgr::loadDAEMatrices (N_DEV_MOSFET1.C:4587) plus the linear
 * part (FilteredMatrix::addToMatrix N_LAS_FilteredMatrix.C:632). nnz each,
 * added to the caller's values as in xg_load_vectors (the masters' "+="
 * through raw matrix pointers, N_DEV_MOSFET1.C:4629-4865). */
int xg_load_matrices(xg_circuit* c, double* dFdx_vals, double* dQdx_vals);

/* a8: OneStep::obtainResidual / obtainJacobian (N_TIA_OneStep.C:219,469)
 * on the vectors/matrices of the last loads (this handle's contributions).
 * q0/q2 = qHistory[0]/[2].  rhs and jac_vals are overwritten. */
int xg_combine(xg_circuit* c, const double* q0, const double* q2, double* rhs, double* jac_vals);

/* The same four calls on DEVICE arrays (pointers into HIP device memory of
 * the handle's device), so a caller that keeps its vectors and matrices on
 * the GPU pays no host copies.  Same layouts and semantics (accumulation
 * included); each call returns when its outputs are complete. */
int xg_update_state_dev(xg_circuit* c, const double* x, const double* sta_cur, const double* sto_cur,
                        double* sta_next, double* sto_next, int* limited);
int xg_load_vectors_dev(xg_circuit* c, double* F, double* Q, double* B, double* dFdxdVp, double* dQdxdVp);
int xg_load_matrices_dev(xg_circuit* c, double* dFdx_vals, double* dQdx_vals);
int xg_combine_dev(xg_circuit* c, const double* q0, const double* q2, double* rhs, double* jac_vals);

/* ---------------- sparse LU (linear side, a9) ---------------- */

/* Symbolic analysis once (AmesosSolver first call: KLU analyse, BTF + fill
 * ordering, N_LAS_AmesosSolver.C:335).  values may be NULL; if given they
 * steer the transversal toward large entries. */
int xg_lu_symbolic(int n, const int* rowptr, const int* colind, const double* values, xg_lu** out);
/* The same with the factorisation method chosen: 0 automatic (multifrontal
 * when the level-scheduled program would exceed 2e7 flops, 65535 unknowns or
 * 2^24 factor entries), 1 level-scheduled entry LU (small systems), 2
 * multifrontal (supernodal fronts, level by level over the elimination
 * tree; for power grids and gate arrays). */
int xg_lu_symbolic_ex(int n, const int* rowptr, const int* colind, const double* values, int method, xg_lu** out);
int xg_lu_destroy(xg_lu* lu);
/* Numeric refactor with the symbolic pivot order (KLU NumericFactorization,
 * :363).  A pivot that is zero or non-finite, or a multiplier beyond KLU's
 * threshold (|l| > 1/0.001), makes it choose new pivots from these values and
 * refactor, as Xyce's KLU re-pivots at every factorisation (repivot_,
 * N_LAS_AmesosSolver.C:87,318); returns >0 (pivot+1) only if the new pivots
 * meet a zero pivot too (a singular matrix).  If the new pivots still leave
 * multipliers beyond the threshold, those factors are kept and the solves
 * refine x against A in FP64 while the residual falls (at most ir_max
 * corrections, to ir_tol); from then on the handle keeps such factors
 * without choosing pivots again. */
int xg_lu_numeric(xg_lu* lu, const double* vals);
/* Solve A x = b with the last factors (:396); NaN -> XG_E_NAN, x zeroed.
 * transpose != 0 solves A^T x = b (doSolve(reuse_factors, transpose),
 * N_LAS_Solver.h:95-107; Amesos SetUseTranspose): level-scheduled method
 * with FP64 factors only (XG_E_ARG otherwise; Xyce's transient path does not
 * use it). */
int xg_lu_solve(xg_lu* lu, const double* b, double* x, int transpose);
/* xg_lu_numeric / xg_lu_solve on device arrays (HIP device memory); same
 * semantics; each call returns when its outputs are complete. */
int xg_lu_numeric_dev(xg_lu* lu, const double* vals);
int xg_lu_solve_dev(xg_lu* lu, const double* b, double* x, int transpose);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): xg_lu_solve_dev with the last FP64 factors, allocating
 * nothing.  b and x are device arrays of n entries.  The solve's work
 * vectors and NaN flag live in one static 32 MiB work area per GPU (4 Mi
 * doubles; the level-scheduled method needs n + 32 of them, the multifrontal
 * one its front vectors plus 2n + 32); a system that needs more returns
 * XG_E_ARG before any device work, as does a handle whose current factors are
 * refined by its solves (FP32 factors, xg_lu_set_precision 1, or factors with
 * multipliers beyond KLU's threshold).  Every
 * call on one GPU uses the same work area: calls that overlap on a GPU race.
 * NaN -> XG_E_NAN with x zeroed, as xg_lu_solve. */
int xg_lu_solve_noalloc(xg_lu* lu, const double* b, double* x);
/* info[0..5] = factor entries (nnz(L+U), or front elements when
 * multifrontal), factor levels, solve levels, flops per factorization,
 * method (1 level-scheduled, 2 multifrontal), supernodes */
int xg_lu_info(const xg_lu* lu, double* info);
/* Factor precision (the mixed-precision form of AmesosSolver::doSolve):
 * 0 = FP64 factors (default); 1 = FP32 factors, FP32 triangular solves and
 * FP64 iterative refinement x += A^-1_32 (b - A x) until
 * max|b - A x| <= ir_tol * max|b|, at most ir_max corrections.  If the
 * residual stops decreasing, ir_max corrections do not reach ir_tol, or the
 * FP32 pivots fail the threshold test, the same call finishes with FP64
 * factors, so a solve is never less accurate than the FP64 path. */
int xg_lu_set_precision(xg_lu* lu, int precision, double ir_tol, int ir_max);
/* stats[0..6] = corrections of the last solve, its relative residual,
 * solves refined, corrections in total, solves finished by FP64 factors
 * after the refinement stopped, refactors whose FP32 pivots failed the
 * threshold test where the FP64 ones passed (those solve with FP64
 * factors), current precision. */
int xg_lu_precision_stats(const xg_lu* lu, double* stats);

/* ---------------- one circuit's transient, resident on the GPU ---------------- */

/* DC operating point and .TRAN of ONE circuit (a power grid, a gate array)
 * with every vector of the loop in device memory: Transient::doLoopProcess
 * (src/AnalysisPKG/N_ANP_Transient.C:1184) with OneStep's trapezoid, predictor
 * and step control (N_TIA_OneStep.C:135,1946,2124), DampedNewton
 * (N_NLS_DampedNewton.C:362-509) and the DC strategy of
 * Interface::spiceStrategy (N_NLS_NOX_Interface.C:321-437: Newton from
 * initJct, then gmin stepping, then source stepping).  The device loads are
 * this library's circuit, the linear solves its LU (lu_method as in
 * xg_lu_symbolic_ex).  The host takes the decisions from a few scalars read
 * back per Newton iteration. */
typedef struct xg_tran xg_tran;
int xg_tran_create(const xg_circuit_desc* desc, double tstop, double tstep, double hmax, int lu_method,
                   xg_tran** out);
int xg_tran_destroy(xg_tran* t);
/* The LU's factor precision, as xg_lu_set_precision. */
int xg_tran_set_precision(xg_tran* t, int precision, double ir_tol, int ir_max);
/* TIA RELTOL/ABSTOL (local error) and NONLIN-TRAN RELTOL/ABSTOL (Newton
 * weights); defaults 1e-3, 1e-6, 1e-2, 1e-6.  Before the first run only. */
int xg_tran_set_tolerances(xg_tran* t, double reltol, double abstol, double nl_reltol, double nl_abstol);
/* Record, at the DC point and every accepted step (up to record_max points),
 * the k products w_r . x (w: k x n row-major host array; compensated sums)
 * followed by the n_sel unknowns x[sel[i]], or with k = n_sel = 0 the whole
 * solution.  Before the first run only. */
int xg_tran_set_record(xg_tran* t, int k, const double* w, int n_sel, const int* sel, int record_max);
/* Runs (the DC operating point first) until t_until or max_steps attempted
 * steps or the end of the window.  stats[0..13] = accepted steps, rejected
 * steps, Newton iterations, loads, time reached, status (0 ok, -100 no DC
 * point, -101 15 rejections in a row, -102 step below 1e-18 s), finished,
 * DC Newton iterations, DC stage (0 Newton, 1 gmin, 2 source stepping,
 * -1 none: UIC),
 * seconds, singular factorisations, LU refinement corrections, LU solves
 * finished by FP64 factors, points recorded. */
int xg_tran_run(xg_tran* t, double t_until, long long max_steps, double* stats);
/* on != 0: .TRAN ... UIC (NOOP; N_ANP_Transient.C:879-1000): no operating
 * point, the run starts from x = 0 with the devices' state of one transient
 * load there.  Before the first run only. */
int xg_tran_set_uic(xg_tran* t, int on);
/* on != 0: one line on stderr per Newton solve of the DC operating point
 * (the continuation stages of a large circuit can take minutes). */
int xg_tran_set_progress(xg_tran* t, int on);
/* Recorded points: t[max], the integration order of the step that ended
 * there (0 = DC point), y[max * (k + n_sel, or n)]; returns the count copied. */
int xg_tran_points(xg_tran* t, int max_points, double* times, int* order, double* y);
/* The last accepted solution (n entries). */
int xg_tran_solution(xg_tran* t, double* x);

/* ---------------- batched Monte Carlo engine ---------------- */

/* Builds a batch of independent circuits (each evaluated, stamped, factored
 * and solved by one workgroup), one launch per Newton iteration of every
 * instance, with per-instance DCOP + OneStep step control on the GPU.
 * .TRAN 0 tstop with step tstep (both > 0) and maximum step hmax (0: none);
 * record_max >= 0 accepted points recorded per instance (0: none). */
int xg_batch_create(int n_inst, const xg_circuit_desc* descs, double tstop, double tstep, double hmax,
                    int record_max, xg_batch** out);
int xg_batch_destroy(xg_batch* b);
/* Launches `launches` batched Newton iterations (captured in a hipGraph of
 * `graph_chunk` launches when graph_chunk > 0).  Returns when done. */
int xg_batch_advance(xg_batch* b, int launches, int graph_chunk);
/* counters[0..8] = Newton iterations (solves), residual loads, accepted
 * steps, rejected steps, instances finished, instances failed, launches,
 * device time of the last xg_batch_advance (s, HIP events), solves that
 * used the partial-pivoting path (DC iterations and threshold failures) */
int xg_batch_counters(xg_batch* b, double* counters);
/* Per-instance state: time, h, order, mode (0 DCOP, 1 tran, 2 done, <0 failed). */
int xg_batch_instance(xg_batch* b, int inst, double* tstate, double* x);
/* Recorded accepted points of one instance: t[max], x[max*n] (both
 * required, max >= 0); returns count. */
int xg_batch_points(xg_batch* b, int inst, int max_points, double* t, double* x);
/* Launch form.  block: workgroup size (64/128/256/512, 0 = keep).
 * split_eval 1 (default): each launch runs a device-evaluation kernel over
 * all MOSFETs of the iterating instances (one thread per device) and then
 * the per-instance kernel, one Newton iteration per launch; 0: the
 * evaluation is fused into the per-instance kernel, which then runs
 * iters_per_launch iterations (at most 1024) back to back with the
 * instance's state and program kept in LDS (0 = keep); 2: fused, and each
 * group's term arrays
 * (the device contributions of one iteration) kept in LDS too where they fit;
 * -1 = keep.  concurrent_groups 1: the
 * instances' LDS-sized groups run side by side on streams of their own, each
 * group its launches of one xg_batch_advance (of one captured graph) back to
 * back, the groups joined at the end (not after every launch, unless the
 * split form or the learned predictor needs all groups between launches);
 * 0: one after another, -1 = keep.  The trajectories are the same. */
int xg_batch_config(xg_batch* b, int block, int iters_per_launch, int split_eval, int concurrent_groups);
/* Time-integration method of every instance: 0 = trapezoid (OneStep, the
 * default), 1 = Gear12 (variable-step BDF1/BDF2), as .OPTIONS TIMEINT
 * METHOD=TRAP|GEAR selects in the reference (src/AnalysisPKG/N_ANP_Transient.C:384-387;
 * src/TimeIntegrationPKG/N_TIA_Gear12.C replaces N_TIA_OneStep.C).  Only before
 * the first launch (xg_batch_advance / xg_batch_autotune): XG_E_ARG after. */
int xg_batch_set_method(xg_batch* b, int method);
/* DC operating point strategy (Interface::spiceStrategy,
 * src/NonlinearSolverPKG/N_NLS_NOX_Interface.C:321-437): every instance runs
 * Newton from initJct, then gmin stepping (10^p - 10^-4 S on every
 * node-voltage row, p from 4 to -4), then source stepping (independent
 * sources scaled from 0 to 1).  first_stage 1 or 2 starts at gmin or source
 * stepping, as .OPTIONS NONLIN CONTINUATION=GMIN|SOURCE does; only before the
 * first launch.  xg_batch_dc_stage reports the stage that produced an
 * instance's operating point (0 Newton, 1 gmin, 2 source stepping). */
int xg_batch_dc_strategy(xg_batch* b, int first_stage);
/* Launch group `group` (instances grouped by LDS need): info[0] instances,
 * [1] workgroup size, [2] 1 if the fused form keeps their term arrays in LDS
 * (chosen by xg_batch_autotune), [3] dynamic LDS bytes per workgroup,
 * [4] the bytes it would take with the term arrays in LDS.  Returns the
 * number of groups (info left untouched if group >= it). */
int xg_batch_group_info(xg_batch* b, int group, int* info);
int xg_batch_dc_stage(xg_batch* b, int inst, int* stage);
/* Per-phase device timers of the first-scheduled instance (the counterpart of
 * the reference's Residual / Jacobian / Linear-solve StatTop timers,
 * src/UtilityPKG/N_UTL_Stats.h:481-511): enable = 1 resets and starts,
 * 0 stops, -1 only reads.  phase_us[0..14] accumulated microseconds of:
 * x load, device eval, vector gather + residual, norms, Newton decision,
 * Jacobian gather, LU refactor, pivot threshold test, triangular solves,
 * update, step control; phase_us[15] = iterations timed. */
int xg_batch_profile(xg_batch* b, int enable, double* phase_us);
/* Autotuner: times the launch forms (split / fused with several iterations
 * per launch) x workgroup sizes x instance grouping on the real batch and
 * keeps the fastest (workgroups of 128, 256 or 512 threads); for the fused
 * form with LDS-sized groups it then picks each group's workgroup size (64 to
 * 512 threads) and the iterations per launch (4, 8, 16, 32, 64).
 * result[0..9] = block size, grouping (1 = by LDS need), us per batched
 * iteration, split_eval, iterations per launch, workgroup size of groups
 * 0-3, groups side by side (1) or one after another (0). */
int xg_batch_autotune(xg_batch* b, double* result);

/* Learned predictor in a batch (call before the first xg_batch_advance):
 * mode 0 off; 1 collect training pairs, up to pair_capacity accepted steps
 * per instance; 2 serve `net` (in_dim XG_PRED_FEAT, out_dim XG_PRED_WIN):
 * after each launch the network proposes the Newton initial guess of every
 * instance that starts a step, x = x_base + w * y per unknown, w its Newton
 * weight (base: xg_batch_predictor_options; features: see xg_batch_pairs).  A proposal with a non-finite or
 * |y| > 512 entry is not used, and a step whose Newton solve
 * fails from a proposal is retried from the polynomial predictor, so the
 * step control and tolerances are those of the run without it. */
int xg_batch_predictor(xg_batch* b, int mode, xg_mlp* net, int pair_capacity);
/* No counterpart in the reference.  Options of the learned predictor (before
 * the first xg_batch_advance; a trained network goes with the options it was
 * trained under): base 0 = the network corrects the polynomial predictor,
 * 1 = the cubic through the last four accepted points (both computed in FP64;
 * the targets of collected pairs are taken from the same base); deadzone =
 * corrections |y| <= deadzone Newton weights are not applied (default 0). */
int xg_batch_predictor_options(xg_batch* b, int base, double deadzone);
/* Collected pairs (mode 1), one record per window of XG_PRED_WIN unknowns
 * of an accepted step, in units of each unknown's Newton weight
 * w = 1e-2 max(|x_n|, |x_poly|) + 1e-6 (the transient Newton weight of the
 * first iteration, N_NLS_DampedNewton.C:297-337): features x [XG_PRED_FEAT]
 * bf16 = g((x_n - x_{n-1}) / w), g((x_{n-1} - x_{n-2}) / w),
 * g((x_{n-2} - x_{n-3}) / w), g((x_poly - x_n) / w), g((x_cubic - x_poly) / w),
 * x_n (XG_PRED_WIN each; g(v) = sign(v) log2(1 + |v|), x_cubic the cubic
 * through x_n..x_{n-3} at t_{n+1}), then h_{n+1}/h_n, h_n/h_{n-1},
 * h_{n-1}/h_{n-2}, order, (log10 h_{n+1} + 12) / 4, 1, zeros;
 * target t [XG_PRED_WIN] = (x_{n+1} - x_base) / w clipped to +-256; mask
 * m = 1 where the window row exists.  Copies up to max_records into the device buffers
 * (NULL: count only); *count = records copied (or available). */
int xg_batch_pairs(xg_batch* b, long long max_records, unsigned short* x_dev, float* t_dev, float* m_dev,
                   long long* count);
/* c[0..4] = pairs collected, steps started from a network proposal, such
 * steps whose Newton solve failed (retried from the polynomial predictor),
 * windows, mode. */
int xg_batch_pred_counters(xg_batch* b, double* c);

/* ---------------- learned Newton predictor network ---------------- */
/* No counterpart in the reference.  An MLP (hidden_layers GELU layers of
 * `width`, linear output) with BF16 weights/activations and FP32
 * accumulation on the matrix cores, trained by its own backward and Adam
 * kernels.  It proposes the Newton initial guess that
 * OneStep::obtainPredictor (N_TIA_OneStep.C:135) computes by polynomial
 * extrapolation; the polynomial predictor stays the fallback.  All
 * dimensions and row counts are multiples of 64.  Pointers named *_dev are
 * device pointers; bf16 values are passed as their uint16 bit patterns. */
int xg_mlp_create(int in_dim, int width, int hidden_layers, int out_dim, int max_rows, unsigned long long seed,
                  xg_mlp** out);
int xg_mlp_destroy(xg_mlp* net);
/* FP32 master parameters, per layer W [out][in] then b [out]; set = 1
 * uploads (and resets the Adam moments), 0 downloads; *count = total. */
int xg_mlp_params(xg_mlp* net, int set, float* params, long long* count);
/* y_dev [rows][out_dim] = net(x_dev [rows][in_dim]) */
int xg_mlp_forward(xg_mlp* net, int rows, const unsigned short* x_dev, float* y_dev);
/* Forward and backward of loss = sum(mask * (y - t)^2) / count; leaves the
 * gradients for xg_mlp_adam; copies them to grads (host, may be NULL). */
int xg_mlp_loss_backward(xg_mlp* net, int rows, const unsigned short* x_dev, const float* t_dev,
                         const float* mask_dev, int count, double* loss, float* grads);
int xg_mlp_adam(xg_mlp* net, float lr, float beta1, float beta2, float eps);
int xg_mlp_sync(xg_mlp* net);

/* ---------------- host-only introspection (no GPU needed) ---------------- */

/* The setup xg_circuit_create runs on the host (pattern, node order, stamp
 * slots), returned without creating device state.  Arrays as in
 * xg_circuit_pattern; sizes as in xg_circuit_sizes.  No reference counterpart:
 * for checking the pattern on machines without a GPU. */
int xg_host_pattern(const xg_circuit_desc* desc, int* sizes, int* rowptr, int* colind, int* node_gid, int* mos_lid,
                    int* mos_fslot, int* mos_qslot, int* mos_stasto);
/* The symbolic analysis of xg_lu_symbolic on the host: info[0..3] as the
 * first four entries of xg_lu_info (nnz(L+U), factor levels, solve levels,
 * flops per factorization);
 * B = A(perm_row, perm_col); L+U pattern of B (lu_colind filled if it fits in
 * max_lu entries).  No reference counterpart (KLU's analyse is internal). */
int xg_host_lu_analyse(int n, const int* rowptr, const int* colind, const double* values, double* info,
                       int* perm_row, int* perm_col, int* lu_rowptr, int* lu_colind, int max_lu);
/* The multifrontal plan of xg_lu_symbolic_ex(method 2) on the host:
 * info[0..5] = supernodes, levels, front elements, flops per factorization,
 * largest front order, nnz(L+U).  No reference counterpart. */
int xg_host_mf_plan(int n, const int* rowptr, const int* colind, const double* values, double* info);

#ifdef __cplusplus
}
#endif
#endif
