// smg_device.hpp -- device-side data layout of the relaxation engine (gfx950).
//
// Particles live in structure-of-arrays form in slot order.  Every job owns a fixed
// slot range [p0, p0 + cap) with cap a multiple of the 256-slot tile, so a tile (one
// workgroup) never straddles jobs and the job's parameters are wave-uniform.  Slots
// [p0, p0 + n) hold live particles in cell-sorted order; the tail holds dead slots
// that sort into the job's pad cell.
#pragma once
#include <cstdint>

namespace smg {

constexpr int kTile = 256;       // slots per workgroup
constexpr int kMaxLevels = 8;    // cell-list levels (cell size c0 / 2^L)
constexpr int kJBits = 26;       // neighbour entry = job-relative slot | (mirror code << 26)
constexpr uint32_t kJMask = (1u << kJBits) - 1u;
// A neighbour entry (lists, tile halos, downloads) names a slot of the same job
// relative to the job's first slot p0, with the mirror code above kJBits: one job holds
// at most 2^26 slots, a pool of jobs any number.
constexpr uint32_t nb_pack(int64_t slot, int64_t p0, uint32_t code) {
  return uint32_t(slot - p0) | (code << kJBits);
}
constexpr int64_t nb_slot(uint32_t e, int64_t p0) { return p0 + int64_t(e & kJMask); }
constexpr int kScanItems = 4096; // table entries per scan workgroup (256 threads x 16)
// Neighbour tiles of the pair passes.  At every list build, each 256-slot tile gets a
// halo: the distinct (slot, mirror code) pairs its candidate lists name, the tile's own
// 256 slots first (halo index = slot - tile start), the others after them.  The tile's
// candidate lists are rewritten as items: 16-bit halo indices, one run per slot, all the
// tile's runs back to back.  The pair passes stage the halo's records in LDS once per
// tile and walk the items with the work split evenly over the workgroup's threads.
constexpr int kRecbPad = 4;      // records past the last slot of recb (the list build's 4-wide loads)
constexpr int kHashN = 4096;     // LDS hash slots of the halo build (keys of one pass) ...
constexpr int kHashSmallIps = 256;   // ... for up to this many pair items per slot; 2 kHashN above
constexpr int kHL1 = 1536;       // halo records the stage-1 pass stages in LDS (24 B each)
constexpr int kHL2 = 1280;       // ... and the stage-2 pass (40 B each)
constexpr int kMaxHalo = 65536;  // halo entries of a tile (16-bit items)
constexpr uint32_t kHMir = 1u << 29;   // staged type word: the record is a mirror image
// overflow bits of JobRed::overflow
constexpr int kOverSpill = 1, kOverItems = 2, kOverHalo = 4;
constexpr int kJumpMax = 64;     // particles per job that may outrun their lists between builds
constexpr int kJumpList = 512;   // candidate entries of one jumper's own list
constexpr int kJumpBit = 1 << 30;   // flag in the type bits of a record: the slot is a jumper

// Level-set cells live in three arrays of the same (bricked) cell order: phi in FP64
// (used for projection), float4 (raw normal xyz, interface tag as int bits) and the
// tag again as one byte.  The interior particles' samples need phi only and read 8
// bytes per corner; the tag lookup reads the byte array.
constexpr int64_t kLsCellBytes = 8 + 16 + 1;

// Refined level-set levels above a job's base grid (a multi-resolution extension: the
// reference's level set is single-level).  Level l has num_pkg << l packages per
// axis; only the packages of a band near the surface are stored, as bricks of 4^D
// cells in package order (cell c of package p is global cell 4 p + c), reached
// through a package map of (packages + 2)^D entries (one halo package on each side):
// brick index, or -1 where the level holds no data.
constexpr int kMaxLsLevels = 3;

// Level-set cells are stored in bricks of 4 x 4 (x 4 in 3D) cells, so that the 2^D
// corners of a sample, and the samples of nearby particles, share few cache lines.
// Cell (a, b, c) of a job's one-cell-halo grid of ng[0] x ng[1] x ng[2] cells sits at
// this offset from the job's first cell; bricks run x fastest, cells inside a brick
// too, and the grid is padded to whole bricks.  (constexpr: host and device.)
constexpr int64_t ls_cell_offset(int D, const int32_t* ng, int a, int b, int c) {
  const int64_t nb0 = (ng[0] + 3) >> 2, nb1 = (ng[1] + 3) >> 2;
  const int bz = D > 2 ? 4 : 1;
  const int64_t brick = (int64_t(c >> 2) * nb1 + (b >> 2)) * nb0 + (a >> 2);
  return brick * (16 * bz) + ((((c & 3) << 2) + (b & 3)) << 2) + (a & 3);
}
constexpr int64_t ls_cells_padded(int D, const int32_t* ng) {
  return int64_t((ng[0] + 3) >> 2) * ((ng[1] + 3) >> 2) * (D > 2 ? (ng[2] + 3) >> 2 : 1) * (D > 2 ? 64 : 16);
}

struct JobParams {
  int32_t D, nlev, profile, remap_max;
  int32_t sym[3], ng[3];
  int32_t num_pkg[3];
  int32_t lev_n[kMaxLevels][3];
  int32_t lev_bits[kMaxLevels];
  double box_l[3], box_r[3], domain[3], pbox_l[3], pbox_r[3];
  double dpkg[3], dcell[3], dl;
  double cell_ratio, cut_off, nu_coeff;
  double min_dl, max_dl, max_phi, tanh25, lin_dx0, lin_dx1, lin_band;
  double c0, clen[kMaxLevels], cinv[kMaxLevels];
  double cbox_l[3], cdom[3];                  // the cell grid: origin and extent
  float wl_h[3], wl_l[3], wh_h[3], wh_l[3];   // walls as float pairs
  float cos_thres, kdx;                       // cos(angle_thres), cut_off/cell_ratio
  float pf_tol;                               // FP32 prefilter bound on |dr| error
  float skin_lim;                             // rebuild when max drift / h_b reaches this
  double cut_v;                               // list radius factor (1 + skin) * cut_off
  int64_t spill_base, spill_cap;              // this job's part of the spill pool
  uint64_t seed;
  int64_t lset_off;                           // first cell of this job's halo grid
  int32_t ls_nl;                              // refined level-set levels (0: single-level)
  int32_t lnp[kMaxLsLevels][3];               // packages per axis of refined level l + 1
  double ldpkg[kMaxLsLevels][3], ldcell[kMaxLsLevels][3];
  int64_t lmap_off[kMaxLsLevels];             // the level's package map in Work::lpm
  int64_t lcell_off[kMaxLsLevels];            // first cell of the level's bricks in lsp / lsn
  int64_t lbricks[kMaxLsLevels];              // bricks stored for the level
  int64_t lmap_base;                          // this job's package maps in Work::lpm
  double seed_pad;                            // seeding margin: one package of the finest level
  int64_t tbl_base, tbl_size;                 // cell-table segment (last entry = pad cell)
  int64_t lev_off[kMaxLevels];
  int64_t p0, n, cap;                         // slot range
};

// fixed-point scale of the residual sum: each particle's |v| dt2 / dx is rounded to a
// multiple of 2^-40 before it is added, so the integer sum does not depend on the
// order of the additions (same value alone, in a batch, or after a re-sort)
constexpr double kResScale = 1099511627776.0;   // 2^40

// per-job reductions, double buffered by step parity; doubles stored as their bit
// patterns (all values >= 0) so that atomicMin / atomicMax on the integer order work
// Accumulators that every workgroup of a pass updates once, kept in kAccCopies copies
// per job, each on cache lines of its own: workgroup b adds into copy b % kAccCopies,
// so one launch's end-of-workgroup atomics spread over that many lines instead of
// queueing on one.  The copies are the accumulation targets; readers fold them (min /
// max / sum), which gives the same value in any order.  Field meanings as in JobRed.
constexpr int kAccCopies = 16;
struct alignas(256) JobAcc {
  unsigned long long dt1[2], dt2[2], maxv[2];
  unsigned long long res[2], maxdisp[2], rsum[2];
  unsigned long long rcum, drift;
  int32_t maxnb, pad_;
};
static_assert(sizeof(JobAcc) == 256, "one JobAcc per 256 bytes");

struct JobRed {
  // dt1 / dt2 / maxv of the previous step folded from the JobAcc copies at the start
  // of each step (the integration in k_map reads them); every other field of the same
  // name in JobAcc is read from the copies
  unsigned long long dt1[2], dt2[2], maxv[2];
  unsigned long long res[2];       // max_i |v_i| dt2 / dx_i
  unsigned long long maxdisp[2];   // max_i |v_i| dt2
  unsigned long long rsum[2];      // sum_i |v_i| dt2 / dx_i in units of 2^-40
  unsigned long long rcum;         // rsum summed over all steps so far (windowed means)
  int64_t step;                    // steps taken by this job (re-seeding RNG stream)
  unsigned long long drift;        // max_i (|x_i - x_b| + cut_off |h_i - h_b|) / h_b this step
  unsigned long long spill_used;   // entries of the job's spill part handed out
  unsigned long long items_used;   // items of the job's part of the item pool handed out
  int64_t nrebuild;                // list builds so far
  int32_t nremap[2];
  int32_t fail, overflow;
  int32_t maxnb;                   // host copy only: folded from the JobAcc copies
  int32_t rebuild;                 // this step re-sorts and rebuilds the lists
  int32_t built;                   // the lists match the current slot order
  int32_t njump;                   // jumpers since the last build (jset holds the first kJumpMax)
  int32_t jover;                   // more jumpers than kJumpMax: rebuild
};

// one buffer set of the fields that follow the particles through the sort
struct PState {
  float* xh[3];
  float* xl[3];
  double* h;
  int32_t* type;
  float* nrm[3];
  int32_t* id;
};

// work arrays and tables
struct Work {
  const JobParams* jp;
  const int32_t* tile_job;
  const int32_t* sblk_job;   // job owning each 4096-entry block of the cell tables (-1: none)
  JobRed* red;
  JobAcc* acc;      // kAccCopies per job, [job * kAccCopies + copy]
  const double* lsp;   // level-set phi per cell
  const float4* lsn;   // level-set (normal, tag bits) per cell
  const int8_t* lst;   // the tag alone, one byte per cell: the per-step tag lookup of
                       // every particle reads a 16x smaller array
  const int32_t* lpm;  // refined level-set package maps (brick index or -1)
  float* nu_a;   // nu read by the stage-2 pass (previous step's value)
  float* nu_b;   // nu written by the stage-2 pass
  float* a1[3];
  float* a2[3];
  float* sum;
  uint32_t* key;
  uint32_t* skey;
  int32_t* perm;     // sort: source slot of each sorted position (cell order)
  int32_t* sperm;    // the same with each cell in particle-id order
  int32_t* sid;      // particle id at each sorted position (k_scatter)
  int32_t* counts;
  int32_t* starts;
  int32_t* bsum;
  // neighbour-read record of a slot, 64 bytes (one half cache line) at rec[4 slot]:
  // [0] = (xh.x, xh.y, xh.z, h), [1] = (xl.x, xl.y, xl.z, type bits),
  // [2] = (a1.xyz, nu of the previous step), [3] = (normal.xyz, 0)
  float4* rec;
  // candidate lists (all pairs within (1 + skin) cut_off h at the last build): the
  // first K entries of a slot inline, in groups of four so that one 16-byte load
  // brings four entries -- entry k at [((k / 4) * cap_total + slot) * 4 + k % 4], K a
  // multiple of 4; entries K.. of a longer list in the job's part of the spill pool
  // at spill[nbo[slot] + k - K]
  int32_t* nbl;
  uint32_t* nbnear;   // near bits of the inline entries: bit k & 31 of word k >> 5 of slot s at [(k >> 5) * cap_total + s]
  int32_t* nbc;     // candidate count per slot (inline + spilled)
  int32_t* nbo;     // spill offset per slot (valid when nbc > K)
  int32_t* spill;   // spill pool, spill_cap entries
  int32_t* ovq;     // slots whose lists spill, appended by the list pass
  int32_t* ntrue;   // neighbours within the reference's cutoff (the force passes' count)
  // neighbour tiles (see kHashN): per tile t, halo entries 256.. at halo[t * ht] as
  // (slot | code << 26), hcnt[t] = 256 + their number (0: not built); items of slot s at
  // items[tbase[t] + ioff[s] ..] (nbc[s] of them), tile t's items taken from its job's
  // part of the pool (ips per slot on average); lane[t * 256 + r] = the tile's particle
  // of list-length rank r (the pair passes' thread order)
  uint16_t* items;
  uint16_t* items2;  // per slot, from the same offset: its stage-2 pairs (nbc2 of them)
  uint32_t* halo;
  int32_t* hcnt;
  int64_t* tbase;
  int32_t* ioff;
  int32_t* nbc2;
  uint8_t* lane;      // lane order of the pair passes
  float4* recb;     // position (hi parts) and h at the last build: (x, y, z, h); h < 0 marks a jumper
  int32_t* jset;    // per job, the jumper slots (kJumpMax each), sorted by slot at each step
  int32_t* jl;      // per job and jumper index, that jumper's candidate list (kJumpList entries)
  int32_t* jcnt;    // its length
  unsigned long long* ctr;   // [0] = ovq length
  int64_t spill_cap;
  int64_t ips;       // items per slot (a tile holds 256 * ips)
  int64_t ht;        // halo entries per tile (<= kMaxHalo)
  int64_t tile0;   // first tile of a launch (launches over one job's tiles start there)
  int64_t cap_total;
  int64_t n_tbl;
  int32_t K;
};


}  // namespace smg
