/* C-ABI of the MI355X photon-transport library (libmidsx_amd.so).
 *
 * The reference has no C or FFI layer: its boundary is the C++ class API of
 * include/MIDSX/Core.h, linked as the static library MIDSX::MIDSX.  Each entry
 * point below names the reference interface it stands in for (paths relative
 * to the reference tree).  Plain pointers and sizes only; host buffers stay
 * caller-owned, device copies belong to the context.
 *
 * Errors: every call returns 0 (MDX_OK) or a negative status; the message of
 * the last failure on the calling thread is returned by mdx_last_error().
 * Status classes map the reference's exception types.
 *
 * Threading: one context per device; calls on one context must come from one
 * host thread at a time.
 */
#ifndef MIDSX_AMD_H
#define MIDSX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDX_OK 0
#define MDX_ERR_INVALID (-1) /* std::invalid_argument */
#define MDX_ERR_RANGE (-2)   /* std::out_of_range */
#define MDX_ERR_RUNTIME (-3) /* std::runtime_error, DatabaseException */
#define MDX_ERR_DEVICE (-4)  /* HIP runtime failure */
#define MDX_ERR_NOMEM (-5)   /* device memory / admission budget exceeded */

/* Source kinds: BeamDirectionality, IsotropicDirectionality,
 * RectangularIsotropicDirectionality, DiscIsotropicDirectionality
 * (include/MIDSX/Core/source.h:86-196); Monoenergetic / Polyenergetic
 * spectra (source.h:23-80); PointGeometry (source.h:205-238). */
#define MDX_DIR_BEAM 0
#define MDX_DIR_ISOTROPIC 1
#define MDX_DIR_RECT_ISOTROPIC 2
#define MDX_DIR_DISC_ISOTROPIC 3
#define MDX_SPEC_MONO 0
#define MDX_SPEC_POLY 1

/* Tally kinds: AACuboidVolumeTally (volume_tally.h:59), DiscSurfaceTally
 * (surface_tally.h:58), RectangularSurfaceTally (surface_tally.h:82), and
 * an image plane (a rectangle split into nx x ny pixels recording incident
 * energy; equivalent to nx*ny RectangularSurfaceTally pixels). */
#define MDX_TALLY_AACUBOID 0
#define MDX_TALLY_DISC 1
#define MDX_TALLY_RECT 2
#define MDX_TALLY_IMAGE 3

/* Tally result layout: [quantity][bucket][field], doubles.
 * Volume quantities (volume_quantity.h:13-28): 0 EnergyDeposition,
 * 1 IncidentEnergy, 2 NumberOfPhotons, 3 NumberOfInteractions.
 * Surface quantities (surface_quantity.h:13-28): 0 IncidentEnergy,
 * 1 EntranceCosine, 2 NumberOfPhotons.
 * Buckets: 0 primary, 1 single incoherent, 2 single coherent, 3 multiple.
 * Fields: 0 number of recorded events n, 1 sum of values, 2 sum of squares
 * (VectorValue::getSum / getVariance, quantity.cpp:27-94, follow from these). */
#define MDX_TALLY_NQ 4
#define MDX_TALLY_NBUCKET 4
#define MDX_TALLY_NFIELD 3
#define MDX_TALLY_ACC_SIZE (MDX_TALLY_NQ * MDX_TALLY_NBUCKET * MDX_TALLY_NFIELD)

/* Interpolated table selectors for mdx_interpolate (material_data.h:95-151). */
#define MDX_XS_TOTAL 0
#define MDX_XS_COHERENT 1
#define MDX_XS_INCOHERENT 2
#define MDX_XS_PHOTOELECTRIC 3
#define MDX_SF_INCOHERENT 4
#define MDX_FF_COHERENT 5 /* host tables only */

/* Event counters reported by mdx_run (index into mdx_stats.counters). */
#define MDX_STAT_HISTORIES 0
#define MDX_STAT_STEPS 1
#define MDX_STAT_REAL 2
#define MDX_STAT_RAYLEIGH 3
#define MDX_STAT_COMPTON 4
#define MDX_STAT_PHOTO 5
#define MDX_STAT_KAHN_TRIALS 6
#define MDX_STAT_KAHN_S_TESTS 7
#define MDX_STAT_DEPOSITS 8
#define MDX_STAT_SURFACE_HITS 9
#define MDX_STAT_ESCAPES 10
#define MDX_STAT_NCOUNTERS 12

/* Transport kernels. */
#define MDX_KERNEL_PLAIN 0 /* one history per thread, one simulation per launch */
#define MDX_KERNEL_QUEUE 1 /* persistent lanes refilled from a work counter */

typedef struct mdx_ctx mdx_ctx;

typedef struct {
  int direction;      /* MDX_DIR_* */
  int spectrum;       /* MDX_SPEC_* */
  double position[3]; /* PointGeometry position (cm) */
  double a[3];        /* beam: pass-through point; rect: corner; disc: centre */
  double b[3];        /* rect: edge1; disc: normal */
  double c[3];        /* rect: edge2 */
  double radius;      /* disc radius (cm) */
  double energy;      /* monoenergetic energy (eV) */
  int n_bins;         /* polyenergetic: number of (energy, weight) bins, at most 65535 */
  const double* bin_energy; /* eV */
  const double* bin_weight; /* normalised internally (probability_dist.cpp:34-39) */
} mdx_source;

typedef struct {
  int kind;          /* MDX_TALLY_* */
  double p0[3];      /* cuboid min corner / disc centre / rect and image corner */
  double p1[3];      /* cuboid max corner / rect and image edge1 */
  double p2[3];      /* rect and image edge2 */
  double normal[3];  /* disc normal (rect / image: normalize(edge1 x edge2)) */
  double radius;     /* disc radius */
  int nx, ny;        /* image pixels along edge1 / edge2 */
  int hist_bins;     /* surface tallies: IncidentEnergy histogram bins, 0 = off */
  double hist_lo;    /* eV */
  double hist_width; /* eV */
} mdx_tally;

typedef struct {
  uint64_t counters[MDX_STAT_NCOUNTERS];
  double kernel_ms; /* device time of the transport launch(es), HIP events */
} mdx_stats;

/* Message of the last failed call on this thread ("" if none). */
const char* mdx_last_error(void);

/* One context per device (stands in for the state PhysicsEngine,
 * physics_engine.h:30, and InteractionData, interaction_data.h:37, hold). */
int mdx_create(int device, mdx_ctx** out);
int mdx_destroy(mdx_ctx* ctx);

/* Directory holding epdl/NNN.endf.gz (or full EPDL NNN.txt files) and
 * materials.csv.  Replaces the configure-time MIDSX_DB_PATH
 * (include/MIDSX/Core/config.h.in:3, CMakeLists.txt:65-73). */
int mdx_set_data_dir(mdx_ctx* ctx, const char* dir);

/* InteractionDataHelpers::convertMaterialNameToId / convertMaterialIdToName
 * (include/MIDSX/Core/material_helpers.h:110-121). */
int mdx_material_id(mdx_ctx* ctx, const char* name, int* id);
/* Mass-energy absorption coefficient of dry air (cm^2/g) at n energies (eV),
 * log-log linear between table nodes, as
 * MaterialData::interpolateMassEnergyAbsorptionCoefficient of material 3
 * (material_data.cpp:39-41, derived_quantities.cpp:51-53), whose database is
 * not available: table 0 (default) holds NIST's values for dry air on NIST's
 * energy grid (data/air_muen_nist.csv), the grid the reference's database
 * interpolates; table 1 the AAPM TG-195 values every 0.5 keV
 * (data/air_muen_tg195.csv). */
int mdx_air_muen(mdx_ctx* ctx, int n, const double* energy, double* out);
int mdx_set_air_muen_table(mdx_ctx* ctx, int table);
int mdx_material_name(mdx_ctx* ctx, int id, char* buf, int buflen);

/* InteractionData(material names) (interaction_data.cpp:3-73): builds the
 * FP64 tables of each material, the majorant and the RITA tables, and
 * uploads their FP32 device form. */
int mdx_build_tables(mdx_ctx* ctx, const int* material_ids, int n_materials);
/* ComputationalDomain::getInteractionData (computational_domain.cpp:11-14,
 * 109-123): builds the tables for the background and every grid material. */
int mdx_build_tables_for_domain(mdx_ctx* ctx);

/* ComputationalDomain (computational_domain.cpp:84-107): domain box
 * [0, dim_space] and background material; grids are added in JSON order
 * (first grid containing a point wins).  ids are x-fastest uint8 material
 * IDs (voxel_grid.cpp:119-121, 175-181). */
int mdx_set_domain(mdx_ctx* ctx, const double dim_space[3], int background_id);
int mdx_add_grid(mdx_ctx* ctx, const uint8_t* ids, const int dims[3], const double spacing[3],
                 const double origin[3]);
/* ComputationalDomain(json path) with NIfTI-1 grids (.nii / .nii.gz). */
int mdx_load_scene(mdx_ctx* ctx, const char* json_path);
int mdx_num_grids(mdx_ctx* ctx, int* n);
int mdx_grid_info(mdx_ctx* ctx, int grid, int dims[3], double spacing[3], double origin[3]);
int mdx_get_grid_ids(mdx_ctx* ctx, int grid, uint8_t* out);

/* PhotonSource(spectrum, directionality, geometry) (source.cpp:143-155). */
int mdx_set_source(mdx_ctx* ctx, const mdx_source* src);

/* The tally factories of runSimulation (run_simulation.h:24-27). */
int mdx_set_tallies(mdx_ctx* ctx, const mdx_tally* tallies, int n);
/* Per-voxel energy deposit of every grid (Voxel::dose, voxel.h:12-15,
 * PhysicsEngine::addVoxelDataToComputationalDomain physics_engine.cpp:160-168). */
int mdx_enable_voxel_dose(mdx_ctx* ctx, int enable);
int mdx_reset_tallies(mdx_ctx* ctx);

/* runSimulation(source, engine, surface_init, volume_init, N, run_time)
 * (run_simulation.cpp:6-41): transports histories [first, first + n) with
 * RNG streams keyed by (seed, history).  Blocking.  Tallies accumulate across
 * calls until mdx_reset_tallies (physics_engine.cpp:111-122). */
int mdx_run(mdx_ctx* ctx, uint64_t n_histories, uint64_t seed, uint64_t first_history, int kernel,
            mdx_stats* stats);

/* Launch parameters of the queue kernel: workgroups per compute unit, and
 * the number of lanes of a wave that must be free (resp. parked at a real
 * interaction) before the wave claims new histories (resp. runs the parked
 * interactions, which it also does once as many lanes are parked as can
 * still step).  They change speed, not results; the defaults are 4, 32,
 * 32.  No reference counterpart
 * (the reference's only parallel knob is the OpenMP thread count). */
int mdx_set_queue_params(mdx_ctx* ctx, int blocks_per_cu, int refill_min, int interact_min);
int mdx_get_queue_params(mdx_ctx* ctx, int* blocks_per_cu, int* refill_min, int* interact_min);
/* The event form MDX_KERNEL_QUEUE runs (mdx_run, mdx_run_jobs and their
 * variants): MDX_EVENT_QUEUE, every lane carrying one photon through
 * wave-level refill / step / interaction phases, or MDX_EVENT_BANK, every
 * wave keeping a bank of bank_size photons (96, 128 or 160) in LDS with one
 * list of bank slots per next event (Woodcock step, background flight, real
 * interaction, free) and running each event on up to 64 photons of one
 * list at a time.  Same physics and tallies; the per-block LDS dose mode
 * (MDX_DOSE_F32C_LDS) always runs on MDX_EVENT_QUEUE.  MDX_EVENT_BANK takes
 * history indices below MDX_BANK_MAX_HISTORY (2^42): a launch with first +
 * n above it fails with MDX_ERR_RANGE.  No reference
 * counterpart (the reference runs one history per OpenMP thread,
 * run_simulation.cpp:15-35). */
#define MDX_EVENT_QUEUE 0
#define MDX_EVENT_BANK 1
int mdx_set_event_kernel(mdx_ctx* ctx, int kind, int bank_size);
int mdx_get_event_kernel(mdx_ctx* ctx, int* kind, int* bank_size);
/* Autotuner: times MDX_KERNEL_QUEUE launches of probe_histories on the
 * selected job over a coordinate search of the lane-resident event kernel's
 * three parameters and, with MDX_TUNE_EVENTS, of the bank kernel's (bank
 * size, blocks per CU, refill and interaction batches), keeps the fastest
 * event kernel and parameters (mdx_get_event_kernel, mdx_get_queue_params),
 * and resets that job's tallies; for a job that uses another job's domain
 * (mdx_job_alias_domain), also that domain's voxel and background dose,
 * where the probes deposited.  With MDX_TUNE_ROUNDS each probe runs the job
 * as two slices of one launch, i.e. the multi-job kernels that
 * mdx_run_jobs and graph rounds launch for several jobs (their register and
 * LDS budgets differ from the single-job kernels').  Outputs may be NULL. */
#define MDX_TUNE_ROUNDS 1
#define MDX_TUNE_EVENTS 2
int mdx_autotune(mdx_ctx* ctx, uint64_t probe_histories, int flags, int* blocks_per_cu, int* refill_min,
                 int* interact_min, double* best_ms);

/* Jobs.  A context holds several scenes ("jobs"), each with its own domain,
 * grids, source and tallies, all sharing the context's tables; the per-scene
 * calls above act on the selected job (job 0 exists from mdx_create).  Used
 * to batch many simulations into one launch (no reference counterpart: the
 * reference runs one simulation per runSimulation call). */
int mdx_job_create(mdx_ctx* ctx, int* job);   /* creates and selects */
int mdx_job_select(mdx_ctx* ctx, int job);
int mdx_job_release(mdx_ctx* ctx, int job);   /* frees its device memory */
/* The selected job uses another job's domain (grids and per-voxel dose)
 * instead of its own, e.g. the views of one CT scan.  A job whose domain is
 * in use by another cannot be released before that one. */
int mdx_job_alias_domain(mdx_ctx* ctx, int src_job);
/* Device memory held by one job (grids, dose, tallies) and by the tables. */
int mdx_job_device_bytes(mdx_ctx* ctx, int job, uint64_t* bytes);
int mdx_tables_device_bytes(mdx_ctx* ctx, uint64_t* bytes);
/* One launch of the queue kernel over several jobs' history slices
 * [first[i], first[i] + n[i]) with RNG streams keyed by (seed[i], history):
 * a job's results do not depend on which other jobs share the launch nor on
 * its job slot, so give different jobs different seeds.  Blocking. */
int mdx_run_jobs(mdx_ctx* ctx, int n_jobs, const int* jobs, const uint64_t* first, const uint64_t* n,
                 const uint64_t* seed, mdx_stats* stats);
/* mdx_run_jobs that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory while they run transport
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The launch's job records, prefix offsets, work counters and
 * event counters live in a static device array of the library (about 19 KB
 * per device, held until the process ends) instead of context buffers, so at
 * most 256 jobs per call (more: MDX_ERR_INVALID).  Every job must be set up
 * before (domain, source, tallies; also the tables), and results are the
 * same as mdx_run_jobs's for the same arguments.  Calls that overlap on one
 * device share that array and race: serialise them.  Blocking. */
int mdx_run_jobs_noalloc(mdx_ctx* ctx, int n_jobs, const int* jobs, const uint64_t* first, const uint64_t* n,
                         const uint64_t* seed, mdx_stats* stats);
/* Background free flights for the selected job's queue-kernel runs (default
 * on).  Through the background material outside every grid the photon flies
 * with that material's own cross section instead of the global majorant
 * (physics_engine.cpp:38-44 samples every step with the majorant), up to the
 * next grid's face, where majorant steps resume, or out of the domain, with
 * an overshoot drawn at the majorant as the reference's last step has.  The
 * distribution of real interactions, tally crossings and escapes is the
 * same; only the number of rejected steps changes.  Taken only when the job
 * has no volume (AACuboid) tally, whose per-segment counts (volume_tally.cpp:
 * 5-27) depend on where steps start. */
int mdx_set_flights(mdx_ctx* ctx, int enable);
/* Local majorants for the queue kernel (default on; a domain with one grid).
 * The grid is cut into macro cells of 8 x 8 x 8 voxels; each cell
 * records the material that bounds the total cross section of every
 * material in it at all energies from 1 to 500 keV, and the radius (in
 * cells) around it within which that still holds, and the grid continues
 * unless that material also bounds the background (outside the grid).
 * Inside that radius the Woodcock step samples with the bounding material's
 * cross section instead of the global majorant, and a step that would leave
 * the radius stops at its edge without a collision test (memoryless
 * restart), so the distribution of real interactions is the reference's
 * (physics_engine.cpp:38-66) with fewer rejected steps.  Off: the global
 * majorant everywhere.  No reference counterpart. */
int mdx_set_local_majorants(mdx_ctx* ctx, int enable);
/* The selected job's macro-cell table (cells[3] cells per axis, 0 when there
 * is none; out: class | uniform << 7 | radius << 8 per cell, x fastest, class
 * 0-126 or a low byte of 255 for none, uniform = every voxel of the cell holds
 * the class material, whose voxel IDs the kernel then does not read; may be
 * NULL).  A device context builds the table when the domain is uploaded, a
 * host-only context on this call. */
int mdx_get_macro_cells(mdx_ctx* ctx, int cells[3], uint16_t* out);
/* Where mdx_build_tables builds the Rayleigh RITA tables (FP64 either way):
 * on the device, one wavefront per energy (default for device contexts), or
 * on the host threads (always for host-only contexts).  The adaptive scheme
 * of probability_dist.cpp:62-71,121-175,267-287 in both. */
int mdx_set_rita_builder(mdx_ctx* ctx, int on_device);
/* Voxel-dose accumulation precision of the selected job's grids (it must own
 * its domain): MDX_DOSE_F64 (default, FP64 atomics), MDX_DOSE_F32 (FP32
 * atomics), MDX_DOSE_F32C (FP32 atomics plus a second FP32 that collects the
 * exact rounding error of every add).  The reference keeps FP64 VectorValue
 * sums per voxel (voxel.h:14); mdx_get_voxel_dose returns FP64 in every
 * mode.  Resets the grids' dose. */
#define MDX_DOSE_F64 0
#define MDX_DOSE_F32 1
#define MDX_DOSE_F32C 2
/* MDX_DOSE_F32C_LDS: as MDX_DOSE_F32C, with each block of the single-job
 * queue kernel summing the deposits of grid 0 in LDS (compensated FP32) and
 * adding its block sums to the global compensated sums at its end; grid 0
 * must fit (8 bytes per voxel within the LDS a block can take, about 17 000
 * voxels).  Other launches use the global form. */
#define MDX_DOSE_F32C_LDS 3
int mdx_set_dose_precision(mdx_ctx* ctx, int mode);
/* Run mdx_run_jobs rounds as a replayed hipGraph (round buffer copy, the
 * clears queued by job set-up during the previous round, counter zeroing,
 * queue kernel), at most 1024 jobs per round. */
int mdx_set_graphs(mdx_ctx* ctx, int enable);
/* mdx_run_jobs split in two: the round is queued and the call returns; while
 * it runs, mdx_get_tally, mdx_get_voxel_dose and mdx_get_background_dose copy
 * on a second stream beside it (a job in the running round reads a partial
 * sum), and so does the set-up (domain, grids, source, tallies, voxel dose)
 * of jobs that are not in the round; with graphs on, the clearing of their
 * dose and tally buffers is left to the next round's graph (reading such a
 * buffer before clears it first).  Every other call waits for the round.
 * mdx_wait waits for the round and returns its counters.  One round in
 * flight at a time. */
int mdx_run_jobs_async(mdx_ctx* ctx, int n_jobs, const int* jobs, const uint64_t* first, const uint64_t* n,
                       const uint64_t* seed);
int mdx_wait(mdx_ctx* ctx, mdx_stats* stats);
/* Free and total device memory (hipMemGetInfo), for admission by footprint. */
int mdx_mem_info(mdx_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);
/* A streaming device-to-device copy of `bytes` (read and write once each),
 * timed: the reference kernel for HBM bandwidth and traffic counters. */
int mdx_copy_probe(mdx_ctx* ctx, uint64_t bytes, double* ms);
/* No counterpart in the reference: a measurement probe like mdx_copy_probe.
 * `deposits` scattered voxel deposits into `records` zeroed 32-byte voxel
 * records at hashed indices, each the three device-scope atomics of a
 * transport kernel's deposit (sum, sum of squares, count), timed over the
 * second of two launches: the achievable atomic rate for a dose grid of
 * that size.  counted (may be NULL): the records' deposit counts summed on
 * the host after both launches (2 x deposits), read back at the end. */
int mdx_atomic_probe(mdx_ctx* ctx, uint64_t records, uint64_t deposits, double* ms, uint64_t* counted);
/* Pinned host memory for fast tally and dose reads (hipHostMalloc). */
int mdx_host_alloc(mdx_ctx* ctx, uint64_t bytes, void** out);
int mdx_host_free(void* p);

/* PhysicsEngine::getVolumeQuantityContainers / getSurfaceQuantityContainers
 * (physics_engine.cpp:124-158).  acc: MDX_TALLY_ACC_SIZE doubles; hist:
 * 4 * hist_bins doubles (may be NULL); image: 3 * nx * ny doubles (sum E,
 * sum E^2, count per pixel; may be NULL). */
int mdx_get_tally(mdx_ctx* ctx, int tally, double* acc, double* hist, double* image);
/* VoxelGrid::getVoxel(i).dose (voxel_grid.cpp:43-45, 86-104): per voxel
 * (sum, sum^2) interleaved and the deposit count (64-bit, as the size of the
 * reference's value list); either may be NULL. */
int mdx_get_voxel_dose(mdx_ctx* ctx, int grid, double* sum_sumsq, uint64_t* count);
/* ComputationalDomain background voxel dose: sum, sum^2, count. */
/* The voxel dose of a grid as FP32 pairs (mean deposit per history S/N,
 * relative standard error of that mean) per voxel, x fastest: half the
 * bytes of the FP64 sums, computed on the device.  n_histories = N, the
 * histories the sums hold.  Runs beside a round in flight. */
int mdx_get_voxel_dose_f32(mdx_ctx* ctx, int grid, uint64_t n_histories, float* out);
/* Queues that FP32 conversion on the transport stream (behind a round in
 * flight, before the next), so that a later mdx_get_voxel_dose_f32 with the
 * same grid and N only copies, beside the next round. */
int mdx_stage_voxel_dose_f32(mdx_ctx* ctx, int grid, uint64_t n_histories);
/* Starts the copy of the selected job's staged FP32 dose into out (pinned
 * host memory for a copy that does not block) and returns; out must stay
 * valid until mdx_read_done reports the copy done.  The copy runs on a
 * stream of its own, so that the set-up of other jobs does not wait behind
 * it.  mdx_read_done polls (wait = 0) or waits for the copy of job's dose,
 * then frees the staging; staging again before that is an error. */
int mdx_read_staged_dose_f32_async(mdx_ctx* ctx, float* out);
int mdx_read_done(mdx_ctx* ctx, int job, int wait, int* done);
int mdx_get_background_dose(mdx_ctx* ctx, double out[3]);

/* MaterialData::interpolate*CrossSection / interpolateIncoherentScattering-
 * Function (material_data.h:95-151) and InteractionData::
 * interpolateMaxTotalCrossSection (interaction_data.h:63, material_id < 0),
 * evaluated on the device in FP32 from the packed tables. */
int mdx_interpolate(mdx_ctx* ctx, int material_id, int which, int n, const float* x, float* out);
/* The same functions on the FP64 host tables (the values the device tables
 * are packed from). */
int mdx_interpolate_host(mdx_ctx* ctx, int material_id, int which, int n, const double* x, double* out);
/* FP64 host table nodes: size query with x == NULL. */
int mdx_get_table(mdx_ctx* ctx, int material_id, int which, int* n, double* x, double* y);
/* MaterialData::sampleCoherentScatteringDCS (material_data.h:151) with the
 * uniform given: device FP32 and host FP64. */
int mdx_sample_rayleigh(mdx_ctx* ctx, int material_id, int n, const float* energy, const float* y, float* mu);
int mdx_sample_rayleigh_host(mdx_ctx* ctx, int material_id, int n, const double* energy, const double* y,
                             double* mu);
/* RITA table of one energy row (probability_dist.cpp:121-136): size query
 * with x == NULL; a and b have rows-1 valid entries. */
int mdx_get_rita_row(mdx_ctx* ctx, int material_id, int energy_index, int* rows, double* energy, double* x,
                     double* y, double* a, double* b);
/* IncoherentScattering::interact sampling (photon_interactions.cpp:62-141)
 * on the device: n Compton events at the given energies. */
int mdx_sample_compton(mdx_ctx* ctx, int material_id, int n, const float* energy, uint64_t seed, float* mu,
                       float* energy_out);
/* ComputationalDomain::getVoxel (computational_domain.cpp:23-31) on the
 * device: grid index (-1 background), linear voxel index, material ID. */
int mdx_voxel_lookup(mdx_ctx* ctx, int n, const float* xyz, int* grid, long long* linear, int* material_id);
/* PhysicsEngine::setInteractionType decision (physics_engine.cpp:82-109) on
 * the device: 0 photoelectric, 1 Rayleigh, 2 Compton. */
int mdx_select_interaction(mdx_ctx* ctx, int n, const float* s_coh, const float* s_inc, const float* s_tot,
                           const float* u, int* type);

/* The delta-scatter decision isDeltaScatter (physics_engine.cpp:180-183)
 * as both transport kernels take it on the device: real[i] = 1 when
 * !(u < 1 - s_tot / s_max) in IEEE FP32 and the step was not cut at a local
 * majorant's bound (limited[i] == 0; such a step takes no collision test),
 * else 0. */
int mdx_delta_decision(mdx_ctx* ctx, int n, const float* s_tot, const float* s_max, const float* u,
                       const int* limited, int* real);
/* Region-of-interest dose statistics: for box b of job jobs[b] (its domain's
 * grid 0, FP64 voxel dose), voxel index ranges boxes[6 b ..] = (i0, i1, j0,
 * j1, k0, k1), half-open, the deposits, their sum and their sum of squares
 * over the box, out[3 b ..] = (n, sum, sum^2): the VectorValue of the box's
 * voxels together (quantity.cpp:27-94 then give its sum and sigma).  One
 * launch for all boxes; the serving loop's stopping test.  No reference
 * counterpart (the reference reads voxels one by one, voxel_grid.cpp:86-104). */
int mdx_roi_stats(mdx_ctx* ctx, int n, const int* jobs, const int* boxes, double* out);
/* PhotonSource::generatePhoton (source.cpp:150-155) on the device for the
 * selected job's source: the directions (3 floats per photon) and energies
 * of histories 0 .. n-1 under seed, as the transport kernels draw them. */
int mdx_sample_source(mdx_ctx* ctx, int n, uint64_t seed, float* dir, float* energy);
/* The device's counter-based generators (no reference counterpart: the
 * reference draws from std::mt19937, probability_dist.cpp:4), for known-
 * answer tests: words 4, Philox-4x32-10 of counter[4 i .. 4 i + 3] under
 * key[2 i], key[2 i + 1]; words 2, Philox-2x32-10 of counter[2 i ..] under
 * key[i]; out holds n * words words. */
int mdx_philox(mdx_ctx* ctx, int n, int words, const uint32_t* counter, const uint32_t* key, uint32_t* out);
/* Particle::rotate (particle.cpp:7-45) as the transport kernels apply it: the
 * direction dir[3 i ..] (unit) turned by polar cosine mu[i] and azimuth
 * phi[i] in [0, 2 pi) into out[3 i ..], on the device. */
int mdx_rotate(mdx_ctx* ctx, int n, const float* dir, const float* mu, const float* phi, float* out);

#ifdef __cplusplus
}
#endif

#endif /* MIDSX_AMD_H */
