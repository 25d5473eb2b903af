// Data layout shared by the host table builder, the C-ABI layer and the HIP
// kernels.  Everything the kernels read lives in one of these plain structs;
// host code fills them and copies them to device memory once per upload.
#pragma once
#include <stdint.h>

#define MDX_MAX_MATERIALS 32
#define MDX_MAX_GRIDS 8
#define MDX_MACRO_SHIFT 3  // macro cells of 8^3 voxels (local majorants)
#define MDX_MACRO_EDGE (1 << MDX_MACRO_SHIFT)
#define MDX_MACRO_UNIFORM 0x80u  // the macro cell holds one material, its class
#define MDX_MAX_TALLIES 16

// Tally quantity / bucket layout (mirrors include/MIDSX/Core/volume_quantity.h:13-28,
// surface_quantity.h:13-28 and the four buckets of *_quantity.cpp measure()).
#define MDX_NBUCKET 4  // 0 primary, 1 single incoherent, 2 single coherent, 3 multiple
#define MDX_NFIELD 3   // n, sum, sum of squares (per recorded event)
// Copies of each tally's scalar records: the atomics of the whole GPU on one
// job's few records would otherwise queue on the same cache lines.
#define MDX_ACC_COPIES 16
#define MDX_VQ_ENERGY_DEPOSITION 0
#define MDX_VQ_INCIDENT_ENERGY 1
#define MDX_VQ_NUMBER_OF_PHOTONS 2
#define MDX_VQ_NUMBER_OF_INTERACTIONS 3
#define MDX_SQ_INCIDENT_ENERGY 0
#define MDX_SQ_ENTRANCE_COSINE 1
#define MDX_SQ_NUMBER_OF_PHOTONS 2
#define MDX_NQ 4

// One interpolated curve f(X) on FP32 nodes.  For X in [x_k, x_{k+1}):
//   f(X) = (a_hi + a_lo) * exp(c1 t + c2 t^2 + c3 t^3),  t = ln(X / x_k)
// kind 0: log-log linear, clamped to the end values outside [x_0, x_{n-1}]
//         (interpolators.cpp:37-76);
// kind 1: log-log cubic B-spline span polynomials, extrapolated with the end
//         spans (interpolators.cpp:3-31).
struct alignas(32) mdx_node {
  float a_hi, a_lo, c1, c2;
  float c3, x, pad0, pad1;  // in a bin record: pad0 = x_{k+1}, pad1 = k (int32 bits)
};

struct mdx_curve {
  int32_t n;       // nodes
  int32_t kind;    // 0 linear, 1 spline
  int32_t off;     // first node in the node arrays
  int32_t hoff;    // first entry in the hash array
  int32_t nh;      // hash bins
  float lnx0;      // ln of the hash range start
  float inv_dlnx;  // hash bins per unit ln X
  float pad;
};

// RITA lookups: an energy hash over ln E gives the largest grid index below
// each bin's start, and every row has MDX_RITA_GUIDE guide entries, entry j
// holding the number of the row's CDF values below j / MDX_RITA_GUIDE.
#define MDX_RITA_GUIDE 64

struct mdx_material_dev {
  mdx_curve total, coh, incoh, photo, sf;
  int32_t rita_ne;     // energies in the RITA grid
  int32_t rita_eoff;   // offset into rita_E (and row index for rita_guide; 2 floats each in rita_escale)
  int32_t rita_roff;   // offset into rita_row_off (rita_ne + 1 entries)
  int32_t material_id;
  int32_t rita_hoff;   // first entry of the energy hash in rita_ehash
  int32_t rita_nh;     // its bins
  float rita_lnE0;     // ln of the hash range start
  float rita_hinv;     // hash bins per unit ln E
};

struct alignas(16) mdx_f4 {
  float x, y, z, w;
};

// One node i of a RITA row (PENELOPE inverse-CDF interpolation,
// probability_dist.cpp:90-119) in w = 1 + x, which keeps the backward end's
// precision: w_i = w_hi + w_lo and the CDF value y_i = y_hi + y_lo as FP32
// pairs (the FP64 values to about 2^-48; y_hi is y_i rounded to nearest, so
// |y_lo| <= half its ulp), the interval's parameters a_i, b_i and widths
// dw = w_{i+1} - w_i, d = y_{i+1} - y_i, each rounded once from FP64 (0 for
// the row's last node).  32 bytes: one aligned sector per node.
struct alignas(32) mdx_rita_row {
  float w_hi, w_lo, y_hi, y_lo;
  float a, b, dw, d;
};

struct mdx_tables_dev {
  const float* xs;          // node abscissae (all curves)
  const mdx_node* nodes;    // node records (all curves)
  const mdx_node* bins;     // bracket hash (all curves): per bin, the record of its node k (see build_hash)
  const float* rita_E;      // RITA energies (all materials), rounded down to FP32
  const float* rita_escale; // per energy i: 1 / ln(E_{i+1}/E_i) and ln(rita_E[i] / E_i)
  const int32_t* rita_row_off;  // per energy: first row; one extra entry per material
  const mdx_rita_row* rita_rows;
  const int32_t* rita_ehash;    // energy hash (all materials)
  const uint8_t* rita_guide;    // MDX_RITA_GUIDE entries per energy row (all materials)
  mdx_curve majorant;
  int32_t n_mat;
  int32_t pad[3];
  mdx_material_dev mat[MDX_MAX_MATERIALS];
  uint8_t id_to_slot[256];  // material ID -> slot, 255 = unknown
};

// Voxel-dose accumulation precision.
#define MDX_DOSE_F64 0   // FP64 atomics
#define MDX_DOSE_F32 1   // FP32 atomics
#define MDX_DOSE_F32C 2  // FP32 atomics with the rounding error of each add kept in a second FP32
#define MDX_DOSE_F32C_LDS 3  // as F32C, the sums of deposits first per block in LDS (one grid that fits)
// MDX_DOSE_F64 voxel record, 32 bytes (one aligned sector, so a deposit
// dirties one cache line): sum, sum of squares, deposit count (uint64), zero.
#define MDX_DOSE_REC 4

struct mdx_grid_dev {
  float origin[3];
  float spacing[3];
  float inv_spacing[3];  // fl(1 / spacing), for the division-free voxel index
  float dim[3];
  int32_t n[3];
  const uint8_t* ids;  // x-fastest: i + j*nx + k*nx*ny (voxel_grid.cpp:119-121)
  double* dose;        // MDX_DOSE_F64: per voxel one MDX_DOSE_REC-double record (nullptr = off)
  float* dose32;       // MDX_DOSE_F32: (sum, sum^2); MDX_DOSE_F32C: (sum, its error, sum^2, its error)
  unsigned long long* dose_n;  // FP32 modes: per voxel deposit count (MDX_DOSE_F64 keeps it in the record)
  int32_t dose_mode;   // MDX_DOSE_*
  int32_t pad;
  // Local majorants (nullptr = off): per macro cell of MDX_MACRO_EDGE^3
  // voxels, x fastest, a 16-bit word: bits 0-6 the slot of a material whose
  // total cross section is at least every cross section found within
  // `radius` cells of it (Chebyshev distance; outside the grid the background
  // material), bit 7 set when
  // every voxel of the cell holds that material (MDX_MACRO_UNIFORM), bits
  // 8-15 that radius; a low byte of 255 (no class) or radius 0: the global
  // majorant applies.
  const uint16_t* macro;
  int32_t mcx, mcy;    // macro cells along x and y
  float mcell;         // smallest edge of a macro cell (cm)
  int32_t pad2;
};

struct mdx_domain_dev {
  float dim_space[3];
  int32_t bg_slot;
  int32_t n_grids;
  int32_t pad[3];
  mdx_grid_dev grids[MDX_MAX_GRIDS];
  double* bg_dose;  // background voxel: sum, sum^2, n
};

#define MDX_DIR_BEAM 0
#define MDX_DIR_ISOTROPIC 1
#define MDX_DIR_RECT_ISOTROPIC 2
#define MDX_DIR_DISC_ISOTROPIC 3
#define MDX_SPEC_GUIDE 64  // guide entries of a polyenergetic source's CDF
#define MDX_SPEC_MONO 0
#define MDX_SPEC_POLY 1

struct mdx_source_dev {
  int32_t dir_kind;
  int32_t spec_kind;
  int32_t n_bins;
  int32_t pad;
  float pos[3];
  float a[3];  // beam: pass-through point; rect: corner; disc: centre
  float b[3];  // rect: edge1; disc: u
  float c[3];  // rect: edge2; disc: v
  float radius;
  float energy;
  const float* cdf;       // poly: cumulative probabilities
  const float* energies;  // poly: bin energies (eV)
  uint16_t guide[MDX_SPEC_GUIDE];  // poly: guide[j] = number of cdf values below j / MDX_SPEC_GUIDE
};

#define MDX_TALLY_AACUBOID 0
#define MDX_TALLY_DISC 1
#define MDX_TALLY_RECT 2
#define MDX_TALLY_IMAGE 3

struct mdx_tally_dev {
  int32_t kind;
  int32_t hist_bins;  // surface tallies: IncidentEnergy histogram bins (0 = off)
  int32_t nx, ny;     // image pixels
  float p0[3];        // cuboid min / disc centre / rect and image corner
  float p1[3];        // cuboid max / rect and image edge1
  float p2[3];        // rect and image edge2
  float nrm[3];       // surface normal
  float radius;
  float e1e1, e2e2;
  float hist_lo, hist_inv_w;
  double* acc;        // [MDX_ACC_COPIES][MDX_NQ][MDX_NBUCKET][MDX_NFIELD]: a workgroup adds
                      // into copy blockIdx % MDX_ACC_COPIES, a read sums the copies
  double* hist;       // [MDX_NBUCKET][hist_bins] photon counts
  double* image;      // [ny*nx][3]: sum of incident energy, sum^2, count
};

struct mdx_tallies_dev {
  int32_t n;
  int32_t pad[3];
  mdx_tally_dev t[MDX_MAX_TALLIES];
};

// One job of a queue-kernel launch: its scene and its slice of histories
// [first, first + n) with RNG streams keyed by (seed, stream, history).
struct mdx_job_dev {
  const mdx_domain_dev* D;
  const mdx_source_dev* S;
  const mdx_tallies_dev* TL;
  uint64_t first;
  uint64_t n;
  uint64_t seed;
  uint32_t stream;   // 0 from mdx_run_jobs: a job's histories do not depend on its slot
  uint32_t flights;  // 1: free flights through the background (no volume tallies)
};

// mdx_run_jobs_noalloc: job records, prefix offsets, work counters and event
// counters of one launch live in a static device array (see transport.hip),
// which holds at most this many jobs.
#define MDX_NOALLOC_MAX_JOBS 256
// History indices the bank event kernel (MDX_EVENT_BANK) takes: first + n
// of every slice at most this (a bank slot keeps 42 bits of the index).
#define MDX_BANK_MAX_HISTORY (1ull << 42)

// One box of a job's voxel grid whose FP64 voxel records mdx_roi_stats sums:
// voxel index ranges [lo, hi) per axis of a grid with nx, ny voxels in x, y.
struct mdx_roi_req {
  const double* rec;  // the grid's MDX_DOSE_REC-double records
  int32_t nx, ny;
  int32_t lo[3], hi[3];
};

// Event counters, one int64 each; indexes into the counter array.
#define MDX_C_HIST 0
#define MDX_C_STEP 1
#define MDX_C_REAL 2
#define MDX_C_RAYLEIGH 3
#define MDX_C_COMPTON 4
#define MDX_C_PHOTO 5
#define MDX_C_KAHN_TRIAL 6
#define MDX_C_KAHN_S 7
#define MDX_C_DEPOSIT 8
#define MDX_C_SURFHIT 9
#define MDX_C_ESCAPE 10
#define MDX_NCOUNTER 12
