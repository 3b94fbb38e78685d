/* clb.h -- C ABI of the MI355X path for CompLB3D's per-timestep hot path.
 *
 * Plain C types only: pointers, sizes and doubles. The caller owns every host buffer;
 * the library owns device memory. Every entry point returns CLB_OK (0) or a negative
 * CLB_ERR_* code and never aborts; clb_last_error() gives the message of the last
 * failure on a context. One context per host thread; a context holds one GPU's share
 * of the domain (a single device).
 *
 * Lattice conventions: node index n = x + nx*(y + ny*z) (x fastest); the x faces are
 * the inlet (x = 0) and outlet (x = nx-1) planes, y and z are periodic. Mask codes are
 * uint8 material numbers (reference defaults: 0 solid / no dynamics, 1 bounce-back,
 * 2 pore, src/complab_functions.hh:983-988); clb_set_dynamics maps each code to a
 * dynamics kind and relaxation frequency, the way NSdomainSetup / soluteDomainSetup
 * assign Palabos dynamics per material number (src/complab_functions.hh:245-340).
 */
#ifndef CLB_H
#define CLB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLB_OK 0
#define CLB_ERR_ARG (-1)     /* bad argument (null pointer, size out of range) */
#define CLB_ERR_HIP (-2)     /* HIP runtime failure */
#define CLB_ERR_STATE (-3)   /* call out of order (e.g. stepping before init) */
#define CLB_ERR_NOMEM (-4)   /* device allocation failed */
#define CLB_ERR_UNSUP (-5)   /* configuration outside what the kernels are built for */

#define CLB_PREC_FP64 0      /* FP64 storage and arithmetic */
#define CLB_PREC_FP32 1      /* FP32 storage (deviations from the weights; D3Q19 at the initial density
                                profile of clb_flow_init), FP32 collision */

#define CLB_KIND_NONE 0      /* Palabos NoDynamics */
#define CLB_KIND_BB 1        /* Palabos BounceBack (node-based, full way) */
#define CLB_KIND_FLUID 2     /* IncBGKdynamics (flow) / AdvectionDiffusionBGKdynamics (solutes) */

#define CLB_BC_DIRICHLET 0
#define CLB_BC_NEUMANN 1

#define CLB_KINETICS_NONE 0
#define CLB_KINETICS_PLUGIN 1   /* the defineKinetics.hh compiled into the library */

#define CLB_WARM_REF 0       /* reference rule: log10 of the component totals / component species */
#define CLB_WARM_USER 1      /* caller-provided log activities */
#define CLB_WARM_MLP 2       /* MLP prediction (BF16 MFMA), reference rule outside the training box */

#define CLB_BATCH_AUTO 0     /* batched solve: the polish from caller / MLP starts, the full iteration
                                from the reference rule's (clb_set_batch_method) */
#define CLB_BATCH_FULL 1     /* the full iteration from every start */
#define CLB_BATCH_POLISH 2   /* the polish from every start */

#define CLB_SPEC_FUSED 0     /* speciation inside the transport kernel (clb_set_speciation_mode) */
#define CLB_SPEC_SEPARATE 1  /* speciation in its own kernel after transport */

#define CLB_MAX_SOLUTES 8
#define CLB_MAX_BIOMASS 4
#define CLB_MAX_COMPONENTS 8
#define CLB_MAX_SPECIES 32
#define CLB_MAX_PORE_CODES 8

typedef struct clb_ctx clb_ctx;

typedef struct clb_config {
  int device;     /* HIP device ordinal */
  int precision;  /* CLB_PREC_* */
  int nx, ny, nz; /* lattice size including the two x boundary planes; nx >= 4 */
} clb_config;

typedef struct clb_perf {
  double flow_ms;        /* summed device time of D3Q19 step kernels (timing enabled) */
  long long flow_launches;
  double ade_ms;         /* D3Q7 transport + kinetics kernels */
  long long ade_launches;
  double spec_ms;        /* speciation kernels */
  long long spec_launches;
  double mlp_ms;         /* MLP inference kernels */
  long long mlp_launches;
  long long fluid_nodes; /* fluid nodes of the current mask */
  long long reactive_nodes;
  double srg_ms;         /* surrogate metabolic processor kernels */
  long long srg_launches;
  long long stragglers;  /* voxels the fused speciation handed to the persistent kernel */
} clb_perf;

/* ---- context (replaces the MultiBlockLattice3D objects main owns, src/complab.cpp:592, 748-775) */
int clb_create(const clb_config* cfg, clb_ctx** out);
int clb_destroy(clb_ctx* ctx);
const char* clb_last_error(const clb_ctx* ctx);
int clb_sync(clb_ctx* ctx);
int clb_device_name(char* buf, int len);

/* ---- geometry (readGeometry + NSdomainSetup material assignment, src/complab_functions.hh:127-294) */
int clb_upload_mask(clb_ctx* ctx, const uint8_t* mask);          /* N codes */
int clb_download_mask(clb_ctx* ctx, uint8_t* mask);
int clb_set_dynamics(clb_ctx* ctx, const int* kind /*256*/, const double* flow_omega /*256*/);
/* The solute lattices' own dynamics per mask code (soluteDomainSetup,
 * src/complab_functions.hh:296-340): there the bounce-back code is BounceBack, the
 * no-dynamics code NoDynamics, and pore and every biofilm code AdvectionDiffusionBGK --
 * also a biofilm that is impermeable to the flow (permRatio = 0 makes it BounceBack only on
 * the flow lattice, :271-283). Until this is called the solute lattices follow
 * clb_set_dynamics' table; afterwards each lattice keeps its own. Transport, kinetics,
 * speciation, dissolution and the surrogate processor act on the voxels this table makes
 * CLB_KIND_FLUID. */
int clb_set_solute_dynamics(clb_ctx* ctx, const int* kind /*256*/);

/* ---- D3Q19 flow (NSdomainSetup + collideAndStream loops, src/complab_functions.hh:245-294,
 *      src/complab.cpp:597-667). Density rho_in on x = 0, rho_out on x = nx-1. */
int clb_flow_init(clb_ctx* ctx, double rho_in, double rho_out);
int clb_flow_steps(clb_ctx* ctx, int nsteps, double* energy /* nsteps values or NULL */);
/* The driver's flow solves with the convergence tracer on the device: the collideAndStream
 * loops of src/complab.cpp:607-620 (ns_convg1), 1794-1800 and 1836-1840 (ns_convg2), each
 * step's stored average energy into Palabos' util::ValueTracer(u, L, epsilon) (a window of
 * (int)(L/u/2) values, converged once full and |stddev/mean| < epsilon, stddev with n - 1, or
 * a NaN ratio), restated on the device. clb_flow_tracer starts a tracer with an empty window;
 * clb_flow_resolve runs steps, at least one, until it has converged or max_steps ran, and the
 * window persists across calls until the next clb_flow_tracer (the driver declares ns_convg2
 * once and never resets it). *steps: steps run; *converged: 1 if the tracer ended the solve;
 * energy: the steps' energies (max_steps values) or NULL. The steps are launched in batches
 * without a host round trip in between (the host loop of clb_flow_steps reads the energy back
 * after every step); the results equal that loop's with the same tracer. */
int clb_flow_tracer(clb_ctx* ctx, double u, double L, double epsilon);
int clb_flow_resolve(clb_ctx* ctx, int max_steps, int* steps, int* converged, double* energy /* or NULL */);
int clb_flow_download_velocity(clb_ctx* ctx, double* u /*3N, interleaved*/, double* rho /*N or NULL*/);

/* ---- D3Q7 solutes (soluteDomainSetup, src/complab_functions.hh:296-340; omega per
 *      substrate, src/complab.cpp:708-727). omega: S*256 (solute-major, per mask code);
 *      bc_type / bc_value: 2*S (west, east) CLB_BC_* and Dirichlet concentration. */
int clb_set_solutes(clb_ctx* ctx, int S, const double* omega, const int* bc_type, const double* bc_value);
int clb_solutes_init(clb_ctx* ctx, const double* c0 /*S*/);
int clb_couple_velocity(clb_ctx* ctx);                        /* latticeToPassiveAdvDiff, src/complab.cpp:1097-1119 */
int clb_set_velocity(clb_ctx* ctx, const double* u /*3N interleaved*/);
int clb_set_biomass(clb_ctx* ctx, int nb, const double* b /*nb*N*/);
int clb_download_conc(clb_ctx* ctx, double* c /*S*N*/);
int clb_download_biomass(clb_ctx* ctx, double* b /*nb*N*/);

/* ---- reactions inside the step
 * clb_set_kinetics: run_kinetics + update_rxnLattices (src/complab3d_processors_part1.hh:72-127,
 *   345-415) with the compiled defineRxnKinetics; dt = ade_dt; solid_code / bb_code are
 *   the skipped material numbers.
 * clb_set_tableau: EquilibriumChemistry setters + run_equilibrium_biotic
 *   (src/complab3d_processors_part4_eqsolver.hh:1355-1494). S: ns*nc row-major over the
 *   solutes (ns == number of solutes in the coupled step); comp_species[j]: the species
 *   index named like component j, or -1. Solved by the FP64 Newton kernel to the
 *   reference's convergence judge (PCF residual 2-norm < tol).
 * clb_set_mask_update: pore <-> biofilm reclassification on total biomass
 *   (updateLocalMaskNtotalLattices3D, src/complab3d_processors_part2.hh:99-210) with one pore
 *   code and one biofilm code: a pore voxel whose biomass (the fields above 1e-14) reaches
 *   threshold becomes biofilm_code; any other voxel that is neither solid nor bounce-back
 *   becomes pore_code below it. On the planes 1 <= x <= nx-2.
 * clb_set_mask_codes: the same rule with the reference's several material types, after
 *   clb_set_mask_update (which it keeps the switch and threshold of): pore_codes[0..npore)
 *   are the pore types (pore_dynamics; pore_codes[0] the default), bio_codes[i] the biofilm
 *   code of biomass field i (bio_dynamics[i][0]; fields beyond nbio count 0). A pore-type
 *   voxel whose biomass reaches the threshold takes the sum of the codes of the fields
 *   present there (src/complab3d_processors_part2.hh:117-125, 190-198); a voxel of no pore
 *   type drops below it to pore_codes[0], or with several pore types to the type of its first
 *   face neighbour of a pore type in the order +x -x +y -y +z -z, no wrap (:141-186). Every
 *   voxel decides from the mask before the pass (the reference's in-place loop lets a voxel
 *   see the new type of an already visited neighbour; its MPI blocks change which). Every
 *   sum of bio_codes must stay within 1..255. The CA (clb_ca_*) uses the same pore types. */
int clb_set_kinetics(clb_ctx* ctx, int kind, double dt, int solid_code, int bb_code);
int clb_set_tableau(clb_ctx* ctx, int nc, int ns, const double* S, const double* logK, const int* comp_species,
                    double tol, int max_iter);
int clb_set_mask_update(clb_ctx* ctx, int enable, double threshold, int pore_code, int biofilm_code);
int clb_set_mask_codes(clb_ctx* ctx, int npore, const int* pore_codes, int nbio, const int* bio_codes);
int clb_step(clb_ctx* ctx, int nsteps);                       /* collide -> kinetics -> speciation -> stream */
int clb_speciation_stats(clb_ctx* ctx, long long* stats /*3: solves, converged, iterations*/);
/* Where the coupled step solves the speciation (no reference counterpart: an execution
 * choice). CLB_SPEC_FUSED (default): inside the transport kernel, from the concentrations
 * it already holds, with the persistent kernel finishing the few voxels that need more
 * iterations; used when the solutes are the tableau's species and nc <= 4, otherwise the
 * separate path runs. CLB_SPEC_SEPARATE: transport kernel, then the persistent kernel over
 * every reactive voxel. Both meet the same judge. */
int clb_set_speciation_mode(clb_ctx* ctx, int mode);
/* Layout of the solute and flow lattices in the step kernels (no reference counterpart: a
 * memory layout choice). CLB_LAYOUT_COMPACT (default): the steps run on the fluid voxels
 * only, stored contiguously in brick order with a neighbour table (the dense layout moves the
 * cache lines of walls too: in a pack of porosity 0.35 nearly every line holds a fluid
 * voxel). Solutes: used when neither the surrogate processor nor dissolution is on, the
 * speciation (if any) is fused, and the pore <-> biofilm rule keeps every voxel fluid for the
 * solutes. Flow (clb_flow_steps): used unless no-dynamics voxels stream. Otherwise the dense
 * layout runs. The state moves between the layouts on its own; results are the same (the
 * FP32 flow to its store's rounding). CLB_LAYOUT_DENSE: always the dense layout. */
#define CLB_LAYOUT_COMPACT 0
#define CLB_LAYOUT_DENSE 1
int clb_set_layout(clb_ctx* ctx, int layout);
/* Order of the fluid voxels in the compact layout (no reference counterpart: a memory layout
 * choice): bricks of bx x by x bz voxels, in x-fastest order, natural inside each, so that a
 * workgroup's run of voxels covers a compact block of the lattice (bx = nx, by = bz = 1: the
 * natural order; 0: the default). Results do not depend on it. */
int clb_set_compact_order(clb_ctx* ctx, int bx, int by, int bz);
/* Evaluation budget of the fused solve, 2..64 (no reference counterpart: a tuning choice,
 * default 3). A voxel that has not met the judge within it is finished by the persistent
 * kernel in the same step, from the point its fused solve reached; a small budget keeps the
 * transport kernel's wavefronts from waiting on their slowest lane. */
int clb_set_fused_budget(clb_ctx* ctx, int evals);
/* Evaluation budget of the stragglers' first pass, 0..64 (no reference counterpart: a tuning
 * choice, default 16). The first pass continues each straggler's fused solve (Newton on the
 * convex potential with the third-order step) within this budget; the full iteration
 * finishes what it leaves. 0: no first pass, the full iteration takes every straggler.
 * Results meet the same judge either way. */
int clb_set_straggler_budget(clb_ctx* ctx, int evals);
#define CLB_KERNEL_FLOW 0    /* D3Q19 collide + stream (at most 256 threads) */
#define CLB_KERNEL_ADE 1     /* D3Q7 transport without fused speciation (at most 256 threads) */
#define CLB_KERNEL_FUSED 2   /* D3Q7 transport with the fused speciation (at most 512 threads) */
/* Thread-block shape (bx along x, by along y; bx*by a multiple of 64) of one of the step
 * kernels (no reference counterpart: a launch choice; default 64 x 4, fused 64 x 8). */
int clb_set_block_shape(clb_ctx* ctx, int kernel, int bx, int by);

/* Launch-shape autotuner (no reference counterpart; the replacement for Palabos' static
 * block decomposition). Times `steps` (even, >= 2) steps of each configured kernel for every
 * candidate block shape -- and, for the fused speciation, every evaluation budget 2..4 -- on
 * the context's current state, keeps the fastest, and reports it. Every trial starts from
 * the same state, and the state, statistics and counters are restored bit for bit
 * afterwards. Needs device memory for one copy of the state while it runs. */
typedef struct clb_tune_report {
  int flow_bx, flow_by;          /* D3Q19 step (0 when no flow is set up) */
  double flow_ms;                /* its device time per step */
  int ade_bx, ade_by;            /* D3Q7 step without fused speciation (0 when not run) */
  double ade_ms;
  int fused_bx, fused_by;        /* D3Q7 step with fused speciation (0 when not run) */
  int fused_budget;
  double fused_ms;               /* per step, persistent-kernel finish included */
  int trials;
} clb_tune_report;
int clb_autotune(clb_ctx* ctx, int steps, clb_tune_report* report);

/* Immobile solutes (vec_immobile, src/complab.cpp:248-266): neither collided nor streamed,
 * changed only by reactions. immobile: S flags; set before clb_solutes_init. */
int clb_set_immobile(clb_ctx* ctx, const int* immobile);
/* Abiotic kinetics with the compiled defineAbioticRxnKinetics (run_abiotic_kinetics +
 * update_abiotic_rxnLattices, src/complab3d_processors_part1.hh:158-336): mode 0 off,
 * 1 every pore / biofilm voxel, 2 only those with a wall among the 6 face neighbours
 * (surfaceAbioticKinetics3D, src/precipitationVOP.hh:53-145). */
int clb_set_abiotic(clb_ctx* ctx, int mode);
/* Solid-phase ids per voxel (the reference's phase lattice, src/complab.cpp:272-375). */
int clb_set_phase(clb_ctx* ctx, const uint8_t* phase);
int clb_download_phase(clb_ctx* ctx, uint8_t* phase);
/* Precipitation / dissolution node conversion (precipNodeConversion3D,
 * src/precipitationVOP.hh:176-204; dissolNodeConversion3D, src/dissolutionVOP.hh:469-511):
 * voxels that are neither solid_code nor bb_code (clb_set_kinetics) and whose mineral
 * (concentration of mineral_solute) reaches precip_max become precip_code, stamped with
 * phase_id when phase_id > 0; other voxels stamped with phase_id whose mineral falls below
 * reopen become pore_code with phase 0, whatever their mask (a reopened wall voxel restarts
 * at rest). precip_max <= 0 or reopen <= 0 disables that direction. changed: voxels whose
 * mask code changed (nonzero => the caller re-solves the flow). */
int clb_node_conversion(clb_ctx* ctx, int mineral_solute, int pore_code, int precip_code, int phase_id,
                        double precip_max, double reopen, long long* changed);
/* Mineral dissolution at the wetted surface of declared solid phases
 * (surfaceDissolutionKinetics3D, src/dissolutionVOP.hh:229-459) with the compiled
 * defineDissolutionRate hook of the abiotic plugin (defineAbioticKinetics.hh:292): every
 * step, each voxel of the phase lattice (clb_set_phase) whose phase id is declared here
 * dissolves against the averaged water of its open face neighbours; its inventory (the
 * immobile solute substrates[k]) shrinks by at most half per step, and the products go to
 * the open neighbours, split evenly, in the same step. The water is taken after the biotic
 * kinetics; the products join the step's increments before streaming. surface_only = 0
 * lets enclosed phase voxels react too (their products are lost, as in the reference).
 * Reopening is clb_node_conversion's reopen threshold. */
int clb_set_dissolution(clb_ctx* ctx, int enable, int nphase, const int* phase_ids, const int* substrates,
                        int surface_only);
/* One solute's concentration field, at rest (initial_fill of a solid phase, a restart):
 * an immobile solute's amount in every voxel, a mobile solute's at-rest equilibrium in
 * every fluid voxel. */
int clb_set_solute_field(clb_ctx* ctx, int solute, const double* c /*N*/);
/* Surrogate metabolic model (run_surrogate3D, src/complab3d_processors_surrogate.hh:74-295,
 * called at src/complab.cpp:1294-1301): every step, with the biotic kinetics, at every
 * reactive voxel (1 <= x <= nx-2, fluid, not the solid / bounce-back codes of
 * clb_set_kinetics, which also gives dt), each listed microbe holding more than 1e-12 takes
 * up substrates at its Michaelis-Menten estimate and grows at the rate its network predicts
 * (or decays); its increments join the kinetics' before they apply.
 *   n: microbes grown by the surrogate (0 switches it off, at most CLB_MAX_BIOMASS);
 *   ids[n]: each one's biomass field (clb_set_biomass), also the id its network is set by;
 *   Kc[n*S]: half-saturation constants (<half_saturation_constants>);
 *   decay[n]: first-order decay, 1/s (<decay_coefficient>);
 *   vmax[n*S] or NULL: <maximum_uptake_flux>, mmol/gDW/h (0: supply-limited);
 *   max_uptake[n*S] or NULL: <substrate_lower_bounds>, a floor <= 0 (<= -1e29: none);
 *   molar_mass[n] or NULL: <biomass_molar_mass>, gDW/mol (default 24.6);
 *   fix_conc[S], fix_lb[S] or NULL: <fix_concentration>, <fix_lower_bounds>;
 *   use_totals: <fba_concentration_basis> total -- uptake limited by the component totals of
 *   the tableau (clb_set_tableau) and drawn from every species carrying the component.
 * Fields of the metabolic config of src/complab3d_metabolic.hh:394-520. Clears the networks. */
int clb_set_surrogate(clb_ctx* ctx, int n, const int* ids, const double* Kc, const double* decay,
                      const double* vmax, const double* max_uptake, const double* molar_mass,
                      const int* fix_conc, const int* fix_lb, int use_totals);
/* The network of surrogate microbe `id`, in place of the compiled defineSurrogateModel
 * (surrogateModel.hh:252-500): a MATLAB-exported feed-forward net -- inputs mapped by
 * mapminmax, xs = (x - x_offset) * x_gain + x_ymin; tansig hidden layers; one linear output
 * neuron; growth = (y - y_ymin) / y_gain + y_offset (10^that with log_output); growth below
 * 1e-8 or not finite is 0.
 *   n_in inputs, input k = Fin[id][in_sub[k]] (the uptake estimates, mmol/gDW/h);
 *   n_layers weight layers (at most 6) of width[l] neurons (at most 16; the last is 1);
 *   params: per layer W (width[l] x width[l-1], row per neuron; width[-1] = n_in), then b;
 *   closure 0: consumption stays at the estimate (the shipped model); 1: consumption =
 *   growth / yield[s], at most the estimate; 2: growth limited by the scarcest substrate,
 *   min over s of estimate * yield[s], and consumption = growth / yield[s] (the two closures
 *   the reference's model file describes); yield[S] in gDW/mmol, <= 0: not tied to growth.
 * n_layers = 0 removes the network: the microbe then grows at 0 and decays. */
int clb_set_surrogate_net(clb_ctx* ctx, int id, int n_in, const int* in_sub, int n_layers, const int* width,
                          const double* params, const double* x_offset, const double* x_gain, double x_ymin,
                          double y_ymin, double y_gain, double y_offset, int log_output, int closure,
                          const double* yield);
/* Checkpoint (the reference's saveBinaryBlock / loadBinaryBlock checkpoints,
 * src/complab.cpp:1122-1132, 1238-1247, in one file of this library's own format): the
 * evolving state -- mask, phase lattice, flow and solute populations (deferred increments
 * folded in), ADE velocity, biomass, CA age and total lattices -- and the step counters.
 * clb_load_state needs a context set up the same way (sizes, solutes, biomass, precision,
 * and which of flow / solutes / phase lattice / CA exist) and continues bit for bit. */
int clb_save_state(clb_ctx* ctx, const char* path);
int clb_load_state(clb_ctx* ctx, const char* path);
/* Velocity statistics of the flow solve (src/complab.cpp:625-671): stats[0] mean |u| over
 * fluid nodes, [1] max |u|, [2] mean u_x over the plane x = nx-2, [3] fluid nodes counted. */
int clb_flow_stats(clb_ctx* ctx, double* stats /*4*/);
/* stabilizeADElattice3D (src/complab3d_processors_part3.hh:261-310): every fluid node back
 * to the at-rest equilibrium of c0; wall nodes keep their populations. */
int clb_solutes_reset(clb_ctx* ctx, const double* c0 /*S*/);
/* Replay the coupled step from captured HIP graphs of steps_per_graph steps (even). */
int clb_set_graph(clb_ctx* ctx, int enable, int steps_per_graph);

/* ---- CA biofilm: biomass redistribution (the sweep loop of src/complab.cpp:1453-1731)
 * pushExcessBiomass3D / halfPushExcessBiomass3D (src/complab3d_processors_part1.hh:432-894),
 * updateLocalMaskNtotalLattices3D (src/complab3d_processors_part2.hh:99-210) after every
 * sweep, updateAgeDistance3D (src/complab3d_processors_part2.hh:718-842) until the ages
 * settle after the first sweep and every 50th. The reference's sweep depends on its loop
 * order and on std::rand; this one is deterministic: each sweep decides every push from
 * the state at its start, with random neighbour orders drawn from (seed, sweep, voxel),
 * then applies them. The pull step (pullExcessBiomass3D) only carries pushes across MPI
 * subdomain faces and has no counterpart on one GPU. Codes and the pore <-> biofilm
 * threshold are those of clb_set_mask_update; walls are clb_set_kinetics' solid_code and
 * bb_code. Biomass is moved on the planes 1 <= x <= nx-2, with no wrap in y or z. */
#define CLB_CA_DONE 0        /* max total biomass <= bmax + 1e-12 */
#define CLB_CA_STALLED 1     /* 100 sweeps without progress: the excess is boxed in */
#define CLB_CA_SWEEP_CAP 2   /* more than 1000 sweeps */
#define CLB_CA_AGE_CAP 3     /* the age relaxation did not settle in 1000 passes */
#define CLB_CA_NONFINITE 4   /* non-finite biomass: nothing was moved */
typedef struct clb_ca_report {
  int sweeps;              /* push sweeps run */
  int age_passes;          /* age passes run */
  int status;              /* CLB_CA_* */
  long long mask_changes;  /* pore <-> biofilm flips during the call */
  double bmax_before;      /* max total biomass before and after */
  double bmax_after;
} clb_ca_report;
/* Age lattice from the current mask (createAgeDomain3D, src/complab3d_processors_part3.hh:
 * 347-367: -1 wall, 0 pore code, 1 other) and the sign of the distance field
 * (calculateDistanceFromSolid, src/complab_functions.hh:169-230), which stays fixed. */
int clb_ca_init(clb_ctx* ctx);
/* Sweep until the max total biomass is within 1e-12 of bmax (max_bMassRho; with FP32
 * biomass, within 8 FP32 ulps of it), or a stop condition of the reference's loop;
 * half = 1: halfPushExcessBiomass3D. Mask flips are
 * counted in the report (nonzero => the caller updates the flow, as after node conversion). */
int clb_ca_redistribute(clb_ctx* ctx, double bmax, int half, unsigned long long seed, clb_ca_report* report);
/* Age passes alone, until no age changes or max_passes (src/complab.cpp:1739 runs one);
 * passes: passes run, negative when max_passes was reached with ages still changing. */
int clb_ca_update_age(clb_ctx* ctx, double bmax, int max_passes, int* passes);
int clb_ca_download_age(clb_ctx* ctx, int* age /*N*/);
int clb_ca_set_age(clb_ctx* ctx, const int* age /*N*/);
/* The finite-difference biofilm solver's block of the loop (src/complab.cpp:1716-1731):
 * mask and totals (as in the CA), then one explicit step of fdDiffusion3D
 * (src/complab3d_processors_part2.hh:257-430), B += nu * (7-point Laplacian of B) for every
 * biomass field at the voxels of 1 <= x <= nx-2 whose mask code is above 1, with zero flux
 * against codes 0 and 1 and the domain's ends; then mask and totals again. nu: the biomass
 * diffusion number (bioNUinPore of the first microbe, src/complab.cpp:729-736). The report's
 * sweeps is 1; mask_changes counts both updates (nonzero => the caller updates the flow).
 * Needs clb_ca_init (the totals lattice). */
int clb_fd_biomass(clb_ctx* ctx, double nu, clb_ca_report* report);

/* ---- batched speciation (the per-voxel solve of EquilibriumChemistry, standalone)
 * totals: n*nc component totals; var_species[nvar] + logk_var[n*nvar]: per-voxel logK of
 * nvar species (NULL / 0 when the tableau's logK holds everywhere); logc: n*nc log10
 * activities (input for CLB_WARM_USER, output always); iters/resid/conv optional. */
int clb_batch_load(clb_ctx* ctx, long n, const double* totals, int nvar, const int* var_species,
                   const double* logk_var);
int clb_batch_solve(clb_ctx* ctx, int warm, const double* logc_in);
/* Iteration of the batched solve for tableaux of more than four components (the solve of
 * src/complab3d_processors_part4_eqsolver.hh:952-1172 restated; judge and tolerance are the
 * reference's either way). CLB_BATCH_FULL: Levenberg-Marquardt on the log residual, Newton
 * on the convex potential when it stalls. CLB_BATCH_POLISH: Newton on the convex potential
 * with a third-order correction of short steps, the judge evaluated only after short steps,
 * falling back to the full iteration after 10 evaluations; for starts near the solution.
 * CLB_BATCH_AUTO (default): the polish from CLB_WARM_USER / CLB_WARM_MLP starts, the full
 * iteration from the reference rule's. Tableaux of at most four components always use the
 * full iteration. */
int clb_set_batch_method(clb_ctx* ctx, int method);
int clb_batch_fetch(clb_ctx* ctx, double* logc, int* iters, double* resid, uint8_t* conv);
/* The batched solve on caller-owned device buffers, allocating nothing. No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). Same kernels and
 * results as clb_batch_load + clb_batch_solve: d_totals[n*nc] and (nvar > 0)
 * d_logk_var[n*nvar] in device memory; d_logc[n*nc] out (the start for CLB_WARM_USER);
 * d_iters[n], d_resid[n], d_conv[n] optional device outputs (NULL: not written). Needs
 * clb_set_tableau (and for CLB_WARM_MLP a trained, scaled network), set up beforehand.
 * Enqueued on the context's stream, no host synchronisation. Its work-queue head lives in a
 * fixed 256-byte library scratch per GPU that every call on that GPU shares: calls that
 * overlap on one GPU race. n <= CLB_NOALLOC_MAX_SOLVES (the queue head is a 32-bit
 * counter that runs past n by up to the grid's lane count). */
#define CLB_NOALLOC_MAX_SOLVES 1073741824L
int clb_batch_solve_noalloc(clb_ctx* ctx, long n, const double* d_totals, int nvar, const int* var_species,
                            const double* d_logk_var, int warm, double* d_logc, int* d_iters, double* d_resid,
                            uint8_t* d_conv);

/* ---- MLP warm start: din -> H -> H -> dout, tanh, tanh, linear; inputs and outputs
 * min-max scaled to [-1, 1] (surrogate_training/trainSurrogate.py:182-201 convention).
 * theta layout (row-major, out x in): W1[H*din] b1[H] W2[H*H] b2[H] W3[dout*H] b3[dout]. */
int clb_mlp_config(clb_ctx* ctx, int din, int hidden, int dout);
int clb_mlp_set_params(clb_ctx* ctx, const float* theta);
int clb_mlp_get_params(clb_ctx* ctx, float* theta);
int clb_mlp_set_scaling(clb_ctx* ctx, const double* xmin, const double* xmax, const double* ymin,
                        const double* ymax);
int clb_mlp_forward(clb_ctx* ctx, long n, const float* X /*n*din, scaled*/, float* Y /*n*dout, scaled*/,
                    int bf16);
int clb_mlp_grad(clb_ctx* ctx, long n, const float* X, const float* T, float* grad, double* loss);
int clb_mlp_train(clb_ctx* ctx, long n, const float* X, const float* T, int epochs, int batch, double lr,
                  uint64_t seed, double* final_loss);
/* Train the speciation warm start on the loaded batch's own Newton solutions:
 * generates inputs (log10 totals + per-voxel logK) and targets (log activities) on the
 * device, fits the scaling box, trains. */
int clb_mlp_train_from_batch(clb_ctx* ctx, int epochs, int batch, double lr, uint64_t seed, double* final_loss);

/* ---- measurement */
int clb_set_timing(clb_ctx* ctx, int enable);
int clb_get_perf(clb_ctx* ctx, clb_perf* perf);
int clb_reset_perf(clb_ctx* ctx);

/* ---- synthetic workloads (input generation; host code, no GPU needed)
 * Sphere pack: overlapping spheres, radius U[rmin, rmax], centres uniform in the interior
 * x range and periodic in y/z, added until porosity <= target; planes 1 and nx-2 copied
 * to 0 and nx-1 (readGeometry, src/complab_functions.hh:142-155); solid voxels with a
 * pore D3Q19 neighbour get bb_code, other solid solid_code, pore pore_code. */
int clb_gen_sphere_pack(int nx, int ny, int nz, double rmin, double rmax, double porosity, uint64_t seed,
                        int pore_code, int solid_code, int bb_code, uint8_t* mask);
int clb_gen_tableau(int nc, int ns, double logk_min, double logk_max, int smin, int smax, uint64_t seed,
                    double log_t_min, double log_t_max, double* S, double* logK);
int clb_gen_totals_field(int nx, int ny, int nz, int nc, double sigma, double log_t_min, double log_t_max,
                         uint64_t seed, double* totals /*N*nc*/);
int clb_gen_biomass_seed(long N, const uint8_t* mask, int pore_code, double fraction, double bmin, double bmax,
                         uint64_t seed_sites, uint64_t seed_density, double* b /*N*/);

#ifdef __cplusplus
}
#endif
#endif /* CLB_H */
