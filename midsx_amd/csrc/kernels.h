// Launchers of the HIP kernels (transport.hip), called by the C-ABI layer.
#pragma once
#include <hip/hip_runtime_api.h>

#include "mdx_types.h"

namespace mdx_launch {

// Byte offsets of the regions of the static scratch (mdx_run_jobs_noalloc):
// event counters, per-job work counters, prefix offsets, job records; total
// size.  At most MDX_NOALLOC_MAX_JOBS jobs.
struct ScratchLayout {
  size_t counters, work, prefix, jobs, bytes;
};
ScratchLayout scratch_layout();
// The calling thread's current device must be `device`.
hipError_t scratch_address(int device, void** out);

hipError_t transport_plain(const mdx_tables_dev* T, const mdx_domain_dev* D, const mdx_source_dev* S,
                           const mdx_tallies_dev* TL, uint64_t first, uint64_t n, uint64_t seed, uint32_t job,
                           unsigned long long* counters, int grid_blocks, hipStream_t st);
// Queue kernel over the history slices of n_jobs jobs; work[0 .. n_jobs)
// (histories claimed per job) must be zero at launch.  A wave claims new
// histories once refill_min of its lanes are free and runs its parked real
// interactions once interact_min lanes have one (or when nothing else can
// proceed).
hipError_t transport_queue(const mdx_tables_dev* T, const mdx_job_dev* jobs, const uint64_t* prefix, int n_jobs,
                           uint64_t total, unsigned long long* work, unsigned long long* counters, int grid_blocks,
                           int refill_min, int interact_min, uint32_t lds_voxels, hipStream_t st);
// Static LDS of the one-job queue kernel (from the runtime when on_device,
// else an upper bound); MDX_DOSE_F32C_LDS adds 8 B per voxel of grid 0.
size_t transport_queue_static_lds(bool on_device);
// Graph form of the multi-job queue kernel: reads the job count, total,
// prefix and job records from a device round buffer.
hipError_t transport_queue_round(const mdx_tables_dev* T, const char* round, unsigned long long* work,
                                 unsigned long long* counters, int grid_blocks, int refill_min, int interact_min,
                                 hipStream_t st);
// The banked event kernel (photon banks of `bank` slots per wave, 96 or
// 128, in LDS; full-lane batches of one event type): the same contract as
// transport_queue / transport_queue_round.
hipError_t transport_bank(const mdx_tables_dev* T, const mdx_job_dev* jobs, const uint64_t* prefix, int n_jobs,
                          uint64_t total, unsigned long long* work, unsigned long long* counters, int grid_blocks,
                          int refill_min, int interact_min, int bank, hipStream_t st);
hipError_t transport_bank_round(const mdx_tables_dev* T, const char* round, unsigned long long* work,
                                unsigned long long* counters, int grid_blocks, int refill_min, int interact_min,
                                int bank, hipStream_t st);
hipError_t eval_curve(const mdx_tables_dev* T, int slot, int which, int n, const float* x, float* out,
                      hipStream_t st);
hipError_t rita(const mdx_tables_dev* T, int slot, int n, const float* E, const float* y, float* mu,
                hipStream_t st);
hipError_t voxel(const mdx_tables_dev* T, const mdx_domain_dev* D, int n, const float* pos, int* grid,
                 long long* lin, int* mat, hipStream_t st);
hipError_t select(int n, const float* sc, const float* si, const float* stt, const float* u, int* type,
                  hipStream_t st);
hipError_t delta(int n, const float* stt, const float* sm, const float* u, const int* limited, int* real,
                 hipStream_t st);
hipError_t rotate(int n, const float* dir, const float* mu, const float* phi, float* out, hipStream_t st);
// Directions (3 per photon) and energies of the source photons of histories
// [0, n) under seed, as the transport kernels generate them.
hipError_t source(const mdx_source_dev* S, int n, uint64_t seed, float* dir, float* E, hipStream_t st);
// words 4: Philox-4x32-10 of ctr[4 i..] under key[2 i..]; words 2:
// Philox-2x32-10 of ctr[2 i..] under key[i].
hipError_t philox(int n, int words, const uint32_t* ctr, const uint32_t* key, uint32_t* out, hipStream_t st);
hipError_t kahn(const mdx_tables_dev* T, int slot, int n, const float* E, uint64_t seed, float* mu, float* Ep,
                hipStream_t st);
hipError_t zero(double* p, size_t n, hipStream_t st);
hipError_t zero_u32(uint32_t* p, size_t n, hipStream_t st);
hipError_t zero_u64(unsigned long long* p, size_t n, hipStream_t st);
// dst[i] = src[i] for i < n (dst may be pinned host memory).
hipError_t copy_u64(const unsigned long long* src, unsigned long long* dst, int n, hipStream_t st);
// Clears the ranges listed in device memory as {count, (address, bytes) x
// count}; addresses 16-byte aligned, sizes multiples of 4.
hipError_t zero_list(const uint64_t* list, hipStream_t st);
// FP64 RITA tables of the Rayleigh angular distribution, one wave per energy:
// per energy ie, grid xs, CDF ys (count[ie] points, -1 if the cap was hit)
// and parameters as, bs (count - 1) at offset ie * rita_cap().
hipError_t rita_build(const double* lx, const double* ly, int n_ff, double alpha, double inv_mc2,
                      const double* energies, int ne, double* xs, double* ys, double* as, double* bs, int* count,
                      hipStream_t st);
int rita_cap();
// Streaming copy of n16 16-byte words (HBM reference for traffic counters).
hipError_t copy16(const void* src, void* dst, size_t n16, int blocks, hipStream_t st);
// n_dep scattered voxel deposits (three device-scope atomics each) into the
// 32-byte records rec[0 .. n_rec) (atomic-rate reference).
hipError_t atomic_probe(double* rec, uint64_t n_rec, uint64_t n_dep, uint64_t seed, int blocks, hipStream_t st);
// Per requested box (n of them, one block each): (deposits, sum, sum^2) of the
// FP64 voxel records into out[3 i ..].
hipError_t roi_stats(const mdx_roi_req* reqs, int n, double* out, hipStream_t st);
// (mean per history, relative standard error) per voxel from the FP64 voxel
// records (sum, sum^2, count, 0).
hipError_t dose_f32(const double* ss, size_t nv, double n, float* out, hipStream_t st);

}  // namespace mdx_launch
