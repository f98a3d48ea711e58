/* mdc.h -- C ABI of the MI355X force-and-integration engine.
 *
 * Plain pointers and sizes only; every call returns an mdc_status (0 = ok)
 * and leaves a message retrievable with mdc_last_error() (per thread) or
 * mdc_ctx_last_error().  Units are the reference's: nm, ps, amu, kJ/mol.
 * Atom records are the reference's 208-byte data::Atom
 * (mdcraft/data/atom.h:21-38), read and written through a field mask.
 *
 * The entry points replace, at whole-array granularity, the per-atom
 * interfaces the reference dispatches through (file:line in the reference):
 *   neibs::VerletList::update / decide / get      neibs/verlet-list.h:177-235
 *   solver::ISolver::prepare + stages             solver/isolver.h:19-32
 *   solver::Pair::forces / charge / virials       solver/pair.h:28-126
 *   potential::Base force/density/virial          solver/potential/base.h:30-65
 *   potential::BaseTernary prepare/force/virial   solver/potential/base.h:105-131
 *   stepper::Base::make_step                      solver/stepper/base.h:78-82
 */
#ifndef MDC_H
#define MDC_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdc_ctx mdc_ctx;
typedef struct mdc_pot mdc_pot;
typedef struct mdc_sys mdc_sys;
typedef int mdc_status;

enum {
    MDC_OK = 0,
    MDC_EINVAL = 1,  /* bad argument */
    MDC_EHIP = 2,    /* HIP runtime error */
    MDC_ENOMEM = 3,  /* device allocation failed */
    MDC_EFILE = 4,   /* file missing or malformed */
    MDC_ESTATE = 5,  /* call out of order (e.g. step before any force) */
    MDC_EBUDGET = 6, /* request exceeds the memory budget */
    MDC_ERANGE = 7   /* value outside a table or non-finite result */
};

/* precision modes */
enum { MDC_FP64 = 0, MDC_MIXED = 1 };
/* what to compute */
enum { MDC_FORCE = 1, MDC_ENERGY = 2, MDC_VIRIAL = 4 };
/* atom record fields for upload / download */
enum {
    MDC_F_R = 1 << 0, MDC_F_V = 1 << 1, MDC_F_F = 1 << 2, MDC_F_VIR = 1 << 3,
    MDC_F_M = 1 << 4, MDC_F_RCUT = 1 << 5, MDC_F_RNS = 1 << 6, MDC_F_EP = 1 << 7,
    MDC_F_EM = 1 << 8, MDC_F_NC = 1 << 9, MDC_F_BC = 1 << 10, MDC_F_KIND = 1 << 11,
    MDC_F_UID = 1 << 12, MDC_F_TAG = 1 << 13, MDC_F_ALL = (1 << 14) - 1
};

/* ---- context: one device, one ordered stream ---------------------------- */
mdc_status mdc_ctx_create(int device, mdc_ctx** out);
mdc_status mdc_ctx_destroy(mdc_ctx* ctx);
/* How calls on this context wait for the device: 0 blocks (the default),
 * 1 polls the stream, keeping a host core busy but waking sooner after
 * short evaluations.  No reference counterpart. */
mdc_status mdc_ctx_set_spin(mdc_ctx* ctx, int spin);
const char* mdc_ctx_last_error(const mdc_ctx* ctx);
const char* mdc_last_error(void);
int mdc_atom_record_size(void); /* 208 */

/* ---- potentials (immutable after create; may be shared by systems) ------ */
/* LJsCut(a, s, rcut): solver/potential/LJs-cut.cxx:8-20 */
mdc_status mdc_pot_lj_create(mdc_ctx* ctx, double epsilon, double sigma, double rcut, mdc_pot** out);
/* LJs(aVr, rVr, Rcutoff): Lennard-Jones with a polynomial tail that takes
 * V and V' to zero at the cutoff (solver/potential/LJs.cxx:7-256) */
mdc_status mdc_pot_ljs_create(mdc_ctx* ctx, double aVr, double rVr, double rcut, mdc_pot** out);
/* Pair(Rcutoff, Vpar, Nr1, dr): tabulated pair potential, r*phi on nr
 * knots of step dr in reference units (solver/potential/pair.cxx:15-34) */
mdc_status mdc_pot_pair_create(mdc_ctx* ctx, double rcut, const double* rphi, int nr, double dr, mdc_pot** out);
/* Pair(filename): four lines skipped, then "Nr dr rcut" (Angstrom) and Nr
 * values of r*phi (eV*Angstrom), converted as pair.cxx:36-56 does */
mdc_status mdc_pot_pair_create_file(mdc_ctx* ctx, const char* path, mdc_pot** out);
/* EAM(Rcutoff, Femb, Rho, Vpar, Nr1, dr, Nrh1, drh): solver/potential/EAM.cxx:15-53
 * (tables in reference units; rphi is r*phi) */
mdc_status mdc_pot_eam_create(mdc_ctx* ctx, double rcut, const double* F, const double* rho, const double* rphi,
                              int nr, double dr, int nrho, double drho, mdc_pot** out);
/* EAM(filename), single-element setfl with eV/Angstrom conversion:
 * solver/potential/EAM.cxx:55-87 */
mdc_status mdc_pot_eam_create_file(mdc_ctx* ctx, const char* setfl_path, mdc_pot** out);
/* Alloy EAM of 1..4 elements (element e = atom kind e): atom i of element a
 * sums rho_b(r) over neighbours of element b, embeds with F_a, and each
 * pair of elements (a, b) adds r*phi_ab(r)/r.  Tables in reference units,
 * rows concatenated: F [nelements][nrho], rho [nelements][nr], rphi
 * [nelements*(nelements+1)/2][nr] in the setfl pair order (1,1), (2,1),
 * (2,2), (3,1), ...  The reference has no such potential: its EAM reads one
 * element (solver/potential/EAM.cxx:15-87) and its Multi shares one density
 * and one embedding table per pair of kinds (solver/multi.cxx:64-99); with
 * one element this is mdc_pot_eam_create. */
mdc_status mdc_pot_alloy_create(mdc_ctx* ctx, int nelements, double rcut, const double* F, const double* rho,
                                const double* rphi, int nr, double dr, int nrho, double drho, mdc_pot** out);
/* The same from an N-element setfl ("eam/alloy") file, the format of
 * problems/NiAl/Ni.eam.alloy, with the eV/Angstrom conversion of
 * solver/potential/EAM.cxx:55-87 */
mdc_status mdc_pot_alloy_create_file(mdc_ctx* ctx, const char* setfl_path, mdc_pot** out);
/* solver::Multi(nkinds) + add_potential(i, j, pot) for each pair of kinds
 * (solver/multi.cxx:11-151), over single-element EAM objects and tabulated
 * Pair potentials (a Pair member adds no density and no embedding term, and
 * leaves Em, as pair.cxx's virial does): npots entries
 * (kind_i[q], kind_j[q], pots[q]) that together cover every pair; a later
 * entry for the same pair replaces an earlier one, as add_potential does.
 * Atom i's list is split by neighbour kind and the object of (kind_i,
 * kind_j) handles that sub-list: its rho table, its pair term, and F' of
 * its own embedding table at both atoms' densities.  The member objects
 * are copied and may be destroyed afterwards.  At most 4 kinds. */
mdc_status mdc_pot_multi_create(mdc_ctx* ctx, int nkinds, int npots, const int* kind_i, const int* kind_j,
                                const mdc_pot* const* pots, mdc_pot** out);
/* MaiseBPNNPotential(filename, domain): solver/potential/maise_bpnn.cxx:21-30;
 * G2 radial + G4 angular descriptors, one or two species (kind-difference
 * descriptor blocks, tools/bpnn/maise_model.cxx:113-131, and one network per
 * species; an atom's kind selects its network and must be below the number
 * of species), 4-layer tanh networks (limits under mdc_pot_bpnn_describe) */
mdc_status mdc_pot_bpnn_create(mdc_ctx* ctx, const char* maise_path, mdc_pot** out);
/* The model mdc_pot_bpnn_create would build from a MAISE file, read and
 * checked on the host only (no device, no context): arch[4] = (inputs,
 * hidden 1, hidden 2, 1), counts[4] = (G2 terms, G4 terms, distinct G4 eta,
 * distinct (lambda, zeta) pairs), rcut in nm (tools/bpnn/maise_model.cxx:
 * 19-50); G2 and G4 terms count one block.  Any of them may be NULL.  The
 * engine takes 4-layer networks of one or two species with up to 64 inputs
 * and 64 neurons per hidden layer, 16 G2 and 64 G4 terms per block over 8
 * eta and 8 (lambda, zeta) pairs, whole zeta 1..32. */
mdc_status mdc_pot_bpnn_describe(const char* maise_path, int* arch, int* counts, double* rcut);
mdc_status mdc_pot_rcut(const mdc_pot* pot, double* rcut);
mdc_status mdc_pot_destroy(mdc_pot* pot);

/* ---- systems: nsys independent boxes resident in HBM --------------------- */
/* box_lo/box_hi: 3 doubles per system; periodic: 3 flags per system
 * (lattice::Domain, lattice/domain.h:14-126).  periodic_boundary: 1 attaches
 * the reference's boundary::Periodic to the solver (solver/boundary/periodic.h),
 * 0 the null boundary.  Atoms of system s are records
 * [off_s, off_s + natoms[s]) of every upload/download buffer.  natoms[s] may
 * be 0 (an empty system: zero energies in mdc_thermo), also for every s (a
 * batch with no atoms, on which build, compute and step do nothing). */
mdc_status mdc_sys_create(mdc_ctx* ctx, int nsys, const uint64_t* natoms, const double* box_lo, const double* box_hi,
                          const uint8_t* periodic, int periodic_boundary, mdc_sys** out);
mdc_status mdc_sys_destroy(mdc_sys* sys);
/* Uploading positions (MDC_F_R) keeps the list; until its next rebuild,
 * each pair's image is chosen from the positions at every evaluation (the
 * image shifts recorded at the build may no longer hold). */
mdc_status mdc_sys_upload(mdc_sys* sys, const void* atoms, uint32_t fields);
mdc_status mdc_sys_download(mdc_sys* sys, void* atoms, uint32_t fields);
/* Records [first, first + count) of the caller's order into atoms[0, count). */
mdc_status mdc_sys_download_range(mdc_sys* sys, uint64_t first, uint64_t count, void* atoms, uint32_t fields);
/* Switch box `box` off (active = 0) or back on: its atoms are then skipped
 * by every stage exactly as atoms with tag == 0 are (single.cxx:44,
 * verlet.cxx:73), so a finished box of a batch costs no work while the
 * others step.  No reference counterpart: batches of boxes are this
 * engine's. */
mdc_status mdc_sys_set_active(mdc_sys* sys, int box, int active);
/* Slot refill for continuous batching (no reference counterpart): box `box`
 * takes `count` atom records (at most the box's record count fixed at
 * mdc_sys_create), in caller's order [off_box, off_box + count), with a new
 * geometry; the box's remaining records become void: they hold no atom,
 * take part in nothing and download as zero records with tag 0.  Every
 * field comes from the records (forces included: compute before stepping
 * when they are not current).  The other boxes keep their state; the list
 * is rebuilt before the next evaluation.  mdc_sys_set_active covers the
 * box's atoms only. */
mdc_status mdc_sys_refill_box(mdc_sys* sys, int box, uint64_t count, const void* atoms, const double* box_lo,
                              const double* box_hi, const uint8_t* periodic);
mdc_status mdc_sys_natoms(const mdc_sys* sys, uint64_t* n);
/* lanes per atom (1, 2, 4, 8, 16) of the force kernels; changes the list layout */
mdc_status mdc_sys_set_lanes(mdc_sys* sys, int lanes_per_atom);
/* Times 1, 2, 4, 8 and 16 lanes per atom and keeps the fastest, leaving a
 * fresh list and forces for it: for LJ and EAM the kernels of a step
 * (density and force in step mode, max(reps, 20) launches back to back;
 * where the fixed-point step path applies, with the tables read from
 * global memory and staged in LDS, keeping the faster, see
 * mdc_sys_step_path), for BPNN one force evaluation (MDC_FORCE, reps
 * repetitions after a warm-up), at the given precision.
 * us_per_eval (may be NULL) receives the 5 mean times in microseconds.
 * No reference counterpart: the reference has no device launch shape. */
mdc_status mdc_sys_autotune(mdc_sys* sys, const mdc_pot* pot, int precision, int reps, int* best_lanes,
                            double* us_per_eval);
/* Threads per workgroup (64, 128 or 256) of the LJ and EAM kernels.
 * No reference counterpart: the reference has no device launch shape. */
mdc_status mdc_sys_set_block(mdc_sys* sys, int threads);
/* The launch tile of the LJ and EAM kernels: lanes per atom (1, 2, 4, 8,
 * 16) times threads per workgroup (64, 128, 256), timed as
 * mdc_sys_autotune times lanes, keeping the fastest pair.  us (may be
 * NULL) receives the 15 mean times in microseconds, lanes-major.  BPNN
 * potentials are refused (their workgroup shape follows the LDS staging).
 * No reference counterpart. */
mdc_status mdc_sys_autotune_tiles(mdc_sys* sys, const mdc_pot* pot, int precision, int reps, int* best_lanes,
                                  int* best_threads, double* us);
/* Step path of single-element EAM in mixed precision.  A step sequence
 * whose boxes all keep a unique minimum image between rebuilds gathers
 * neighbours as 32-bit fixed-point coordinates (a wrapped integer
 * difference is the minimum image; 16 bytes per neighbour, FP32 pair
 * arithmetic), reading the interval records of the tables from global
 * memory (set = 0) or from a copy staged in LDS by persistent workgroups
 * (set = 2; the default for systems of 65 536 records and more, and
 * mdc_sys_autotune picks between the two); set = 1 keeps the
 * FP64-difference kernels for every step; set = -1 leaves the choice as it
 * is.  *last (may be NULL) receives what the last step sequence or kernel
 * timing of this system ran: 0 the FP64-difference kernels, 1 the
 * fixed-point kernels with global tables, 2 with tables in LDS.
 * No reference counterpart: a choice between device paths. */
mdc_status mdc_sys_step_path(mdc_sys* sys, int set, int* last);
/* Graph mode (on != 0) of mdc_compute on this system: the device work of
 * an evaluation (fills, density and force kernels, the counter readback) is
 * captured in a HIP graph for its configuration (potential, flags,
 * precision, list build, lanes, workgroup size and the arrays it reads)
 * and replayed while the configuration holds; a configuration's first call
 * runs as usual, its second captures it.  Results are those of the plain
 * path.  Evaluations with thermostats attached or with kernel timing on
 * run as usual.  captures and launches (may be NULL) report the counts.
 * No reference counterpart: the reference launches nothing. */
mdc_status mdc_sys_set_graphs(mdc_sys* sys, int on);
mdc_status mdc_sys_graph_stats(const mdc_sys* sys, uint64_t* captures, uint64_t* launches);

/* ---- neighbour list (neibs::VerletList) --------------------------------- */
/* update(kbuf = -1): wraps positions into periodic boxes, bins, rebuilds */
mdc_status mdc_nlist_build(mdc_sys* sys);
/* decide(): 1 when any atom moved more than half the skin (rns - rcut) */
mdc_status mdc_nlist_decide(mdc_sys* sys, double rcut, int* need);
/* update(kbuf > 0): VerletList::optimize_search_radius
 * (neibs/verlet-list.cxx:71-92) over the current list: rns_i = kbuf * max
 * over atom i's list (self included) of rcut_i + rcut_j, with rcut as last
 * uploaded; 0 for atoms of a system with no list yet (the reference's
 * lists start empty).  rns_out (n doubles, caller's order, may be NULL)
 * receives the new radii; the list is out of date until mdc_nlist_build. */
mdc_status mdc_nlist_optimize_rns(mdc_sys* sys, double kbuf, double* rns_out);
/* automatic = 1 (default): step sequences rebuild the list on the half-skin
 * trigger.  automatic = 0: the list is rebuilt only by mdc_nlist_build, as
 * the reference's callers drive VerletList::update themselves
 * (test/python/solver/test-EAM.py: update every NCbuff steps); each pair's
 * image is then chosen from the current positions at every step
 * (solver/boundary/periodic.cxx:36-60).  No reference counterpart: the
 * reference has no automatic rebuild. */
mdc_status mdc_sys_set_rebuild(mdc_sys* sys, int automatic);
/* Attach (1) or detach (0) boundary::Periodic after creation: the boundary
 * argument of solver::Single (solver/single.h, python/mdcraft/solver/_solver.cxx:71-82). */
mdc_status mdc_sys_set_boundary(mdc_sys* sys, int periodic_boundary);
/* full list including self in the reference's form: CSR over original atom
 * order; offsets has n+1 entries; ids NULL queries the size via *total */
mdc_status mdc_nlist_export(mdc_sys* sys, uint64_t* offsets, uint64_t* ids, uint64_t* total);
mdc_status mdc_nlist_stats(mdc_sys* sys, double* mean_neighbours, int* max_neighbours, int* capacity, int* rebuilds);

/* ---- force evaluation and integration ----------------------------------- */
/* prepare + stages (+ virials when MDC_ENERGY | MDC_VIRIAL; for BPNN the
 * virial is the reference's diagonal one, maise_bpnn.cxx:98-120, with the
 * off-diagonal entries zero) */
mdc_status mdc_compute(mdc_sys* sys, const mdc_pot* pot, uint32_t what, int precision);
/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): mdc_compute for a BPNN potential whose
 * work arrays (descriptors, dE/dG, force and virial shares, reverse-slot
 * index; carried by the system and allocated on its first evaluation in
 * mdc_compute) come instead from a fixed 64 MiB work area that the library
 * holds on each device for the life of the process.  Allocates nothing.
 * Needs a built list (MDC_ESTATE otherwise: it builds none) and a BPNN
 * potential (MDC_EINVAL otherwise); an evaluation whose arrays exceed the
 * work area fails with MDC_EINVAL and a message naming the size (13 500
 * atoms with 136 list slots need about 53 MiB in MDC_MIXED with
 * MDC_VIRIAL).  Every call on one device shares that device's work area:
 * calls on one context are ordered by its stream, but calls that overlap on
 * one GPU from different contexts race. */
mdc_status mdc_compute_noalloc(mdc_sys* sys, const mdc_pot* pot, uint32_t what, int precision);
/* Langevin thermostats (solver/thermostat/langevin.cxx:7-42), applied one
 * after the other as thermostat::Compound does, to every force evaluation
 * of the system (mdc_compute, MDC_STAGE_FORCES, each step of mdc_step),
 * after the potential's force (solver/pair.cxx:72-87): per atom whose x
 * lies in [xmin, xmax],
 *   f += ((sigma g) sqrt(m) - (v - U) beta m) * heat,
 *   sigma = sqrt(2 beta kB T / time_step), kB = 0.008314462175 kJ/mol/K,
 * with v the half-step velocity inside a step.  params holds 11 doubles
 * per thermostat in the reference constructor's order: beta, temperature,
 * time_step, heat_x, heat_y, heat_z, Ux, Uy, Uz, xmin, xmax.  count = 0
 * detaches them.  g are normal deviates of a counter-based stream
 * (Philox4x32-10 + Box-Muller) at (seed, atom index, application count,
 * thermostat index): runs are reproducible for a seed, unlike the
 * reference's std::random_device seeding.  A new seed restarts the count.
 * Atoms with tag == 0 get no thermostat force (the reference adds one to
 * their f, though they never move). */
mdc_status mdc_sys_set_thermostat(mdc_sys* sys, int count, const double* params, uint64_t seed);
/* One solver call of the reference's Python API on the resident state
 * (solver::Pair::{prepare,charge,forces,virials}, solver/pair.cxx:16-127;
 * solver::Single, single.cxx:38-83; solver::Ternary, ternary.cxx:19-67):
 *   MDC_STAGE_PREPARE  zero f, V, Ep (and Em, nc for EAM) of atoms with
 *                      tag != 0, set rcut; BPNN: descriptors, network,
 *                      dE/dG and Ep (maise_bpnn.cxx:27-59, :275-325)
 *   MDC_STAGE_CHARGE   EAM density into nc (EAM.cxx:499-560)
 *   MDC_STAGE_FORCES   add the force to f; EAM takes F'(nc) of the nc the
 *                      atoms hold, whether or not a charge call preceded
 *   MDC_STAGE_VIRIALS  add the pair energy and virial to Ep and V, set Em =
 *                      F(nc) for EAM; f is left as it is.  BPNN: adds the
 *                      reference's diagonal virial (maise_bpnn.cxx:98-120)
 *                      from the dE/dG of the last prepare
 * The list must be built. */
enum { MDC_STAGE_PREPARE = 0, MDC_STAGE_CHARGE = 1, MDC_STAGE_FORCES = 2, MDC_STAGE_VIRIALS = 3 };
mdc_status mdc_stage(mdc_sys* sys, const mdc_pot* pot, int stage, int precision);
/* Mean duration in microseconds of the step-mode density (us_out[0], 0 for
 * LJ) and force (us_out[1]) kernels, each launched reps times back to back
 * between two events on the context stream.  Leaves positions, velocities,
 * forces and the list as they were.  LJ and EAM only.
 * No reference counterpart: a measurement hook. */
mdc_status mdc_sys_time_kernels(mdc_sys* sys, const mdc_pot* pot, int precision, double dt, int reps,
                                double* us_out);
/* nsteps of velocity Verlet (make_step), list rebuilt on the device on the
 * half-skin trigger (a list that overflows there is grown and built again
 * before any step uses it); *rebuilds receives the number of rebuilds (may
 * be NULL) */
mdc_status mdc_step(mdc_sys* sys, const mdc_pot* pot, double dt, int nsteps, int precision, int* rebuilds);
/* per system: Ep, Em, KE, Etot, W (sum of virial traces), com velocity[3] */
mdc_status mdc_thermo(mdc_sys* sys, double* out);
mdc_status mdc_sync(mdc_sys* sys);

/* ---- serving: continuous batching of requests (no reference counterpart;
 * the reference runs one system per process) ------------------------------
 * A request is one structure (atom records, box) for one registered
 * potential, with a task: a single-point energy/force/virial evaluation
 * (steps = 0) or `steps` velocity-Verlet steps followed by one.  Requests
 * are admitted in arrival order into free slots of running batches (a
 * batch is one system whose boxes are slots of one size class, 64 *
 * 2^(k/2) records, for one potential and time step), so arrivals join
 * batches that are already stepping; a request that finds no free slot
 * opens a new batch of slots_per_batch slots while the estimated device
 * bytes of all open batches stay within budget_bytes (idle batches are
 * closed first; otherwise the queue waits).  Every batch runs on its own
 * context (stream) and host thread, so batches overlap on the device.
 * Each batch steps in chunks of at most chunk_steps, never past the
 * smallest step count left among its slots; a request whose count is
 * reached gets a final E/F/virial evaluation of the batch and its slot is
 * freed.  submit, wait, take and stats may be called from any thread.  The
 * potentials must outlive the server. */
typedef struct mdc_server mdc_server;
mdc_status mdc_serve_create(int device, double budget_bytes, int chunk_steps, int slots_per_batch, mdc_server** out);
/* register a potential evaluated at `precision`; *kind is its index */
mdc_status mdc_serve_add_potential(mdc_server* srv, const mdc_pot* pot, int precision, int* kind);
/* copies the natoms records and the box (3 doubles each, 3 periodic
 * flags); dt is unused for steps = 0; *id names the request */
mdc_status mdc_serve_submit(mdc_server* srv, int kind, const void* atoms, uint64_t natoms, const double* box_lo,
                            const double* box_hi, const uint8_t* periodic, double dt, int steps, uint64_t* id);
/* waits up to timeout_s (0: no wait) for the request; *state = 1 done,
 * -1 failed, 0 not yet */
mdc_status mdc_serve_wait(mdc_server* srv, uint64_t id, double timeout_s, int* state);
/* a finished request's records after its task (natoms records), its thermo
 * row (Ep, Em, KE, Etot, W, com velocity[3]; may be NULL) and times in
 * seconds since the server started (submitted, admitted, done; may be
 * NULL); releases the request.  A failed request returns its error. */
mdc_status mdc_serve_take(mdc_server* srv, uint64_t id, void* atoms_out, double* thermo, double* times);
/* done, failed, queued, running, open batches, batches opened, open batch
 * bytes, peak open batch bytes */
mdc_status mdc_serve_stats(mdc_server* srv, double* out);
/* finishes the running requests, then stops every batch */
mdc_status mdc_serve_destroy(mdc_server* srv);

/* ---- instrumentation ---------------------------------------------------- */
/* BPNN descriptors G and dE/dG (kJ/mol per descriptor unit) of the last
 * evaluation, [n][nin] in original atom order (the per-atom g and dedg of
 * maise_bpnn.h:40-47); G or dEdG may be NULL; *nin receives the width */
mdc_status mdc_bpnn_inspect(mdc_sys* sys, const mdc_pot* pot, double* G, double* dEdG, int* nin);
/* Measurement hook, no reference counterpart: `reps` launches of a plain
 * device copy of `bytes` (a multiple of 16) from one buffer to another on
 * the device's default stream, after one untimed launch; *us_per_copy (may
 * be NULL) receives the mean time per copy.  Its FETCH_SIZE / WRITE_SIZE
 * counter readings calibrate the engine kernels' traffic figures. */
mdc_status mdc_copy_calibration(int device, uint64_t bytes, int reps, double* us_per_copy);
/* kernel timing with hipEvents on the launch stream: kind 0 pair force,
 * 1 EAM density, 2 pre-drift, 3 BPNN descriptors, 4 BPNN MLP, 5 BPNN force,
 * 6 BPNN gather, 7 rebuild, 8 BPNN fused pass (descriptors, network and
 * force shares in one kernel: FP32, one species).  total_ms and launches
 * hold 9 entries each. */
mdc_status mdc_timing(mdc_sys* sys, int enable, double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif
