// hd_internal.hpp -- shapes and kernel launchers shared between the host
// orchestration (hd_api.cpp) and the kernel sources (*.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <functional>
#include <vector>

namespace hd {

// Geometry of one gulp's DM series (one hd_execute call, scrunch 1).
// Names follow Pipeline/pipeline.cu:488-499 and remove_baseline.cu:21-46.
struct SeriesShape {
  unsigned nbeams;
  size_t stride;       // cur_beam_stride = nsamps/nbeams
  size_t nout;         // nsamps_computed = nsamps - max_delay
  size_t bc;           // cur_beam_nsamps = stride - max_delay
  size_t ppitch;       // plane row pitch in floats (>= nout+1, zero past nout)
  size_t spitch;       // prefix-sum row pitch in floats (>= nout+1)
  // baseline (remove_baseline.cu:38-46)
  size_t sc;           // sample_count
  size_t count_round;  // 5^nscr * sample_count
  int nscr;            // nscrunches
  int bl_direct;       // pyramid levels computed straight from the leaves
  float step_rs_beam;  // float(stride-1)/count_round   (beams 0..nb-2)
  float step_rs_last;  // float(bc-1)/(count_round-1)   (last beam)
  float step_bl_beam;  // float(sc-1)/stride            (beams 0..nb-2)
  float step_bl_last;  // float(sc-1)/(bc-1)            (last beam)
  // per-beam RMS (get_rms.cu:66-100)
  int rms_direct;
  float rms_delta;     // float(stride - bc)
};

// One boxcar width of the filter loop (pipeline.cu:657-692).
struct WidthInfo {
  unsigned w;          // rel_filter_width
  unsigned ts;         // rel_tscrunch_width
  unsigned ahead;      // (w-1)/2+1
  unsigned behind;     // w/2
  unsigned n_w;        // cur_nsamps_filtered
  unsigned filter_idx; // log2 w
  unsigned merge_dist; // cand_sep_time * (w/ts)
  unsigned rel_off;    // (boxcar_max/2)/ts
  unsigned cs;         // cur_scrunch = scrunch factor * ts: output index to sample index
};

constexpr int kMaxWidths = 24;

inline size_t pow5_host(int L) {
  size_t p = 1;
  for (int i = 0; i < L; ++i) p *= 5;
  return p;
}

struct WidthTable {
  int nw;
  unsigned off;        // boxcar_max/2
  WidthInfo wi[kMaxWidths];
};

// Block-RMS pyramid geometry (get_rms.cu:104-131) for one width.
struct BlockRmsShape {
  size_t n;            // pooled count = nbeams * bc692
  size_t bc;           // bc692 = stride - (max_delay + boxcar_max)
  float delta;         // float(stride - bc692)
  int L0;              // levels computed per chunk
  size_t chunk;        // 5^L0
  size_t np;           // n / 5^L0 partial values
};

// One gulp of a batched call: its own input and sample count (the shape of a
// separate hd_execute call), sharing the pipeline's plan (trial list,
// channels, boxcar widths).  An array of these lives in device memory; the
// kernels of a batched call index it with their group coordinate.  Rows of
// the DM plane and of the prefix sums are (group, trial) pairs, row
// g*n_dm + m, all with the same pitch.
struct GroupDev {
  SeriesShape sh;
  BlockRmsShape bs;
  size_t bc692;               // stride - (max_delay + boxcar_max)
  const uint8_t* fb;          // input bytes in device memory
  size_t nsamps;              // input time samples (all STF beams of the group)
  unsigned n_w[kMaxWidths];   // fresh boxcar outputs per width
  uint64_t t_off;             // clustering coordinate of the group's sample 0
  uint64_t nsamps_beam;       // nsamps / nbeams
  // A group may hold a slice of a gulp's trials (the gated search): its rows
  // are trials trial0 .. trial0 + n_rows - 1 of the plan (trial0 a multiple
  // of the dedispersion block), rows >= rows_valid are padding and produce
  // no giants, its unpacked input is slot xslot of x, and its giants belong
  // to clustering group cgroup (the gulp).
  uint32_t trial0 = 0;
  uint32_t rows_valid = 0xFFFFFFFFu;
  uint32_t xslot = 0;
  uint32_t cgroup = 0;
  // dedispersion with two staging buffers for this group (when the launch
  // chooses per group, DedispGeom::dbuf_groups)
  uint32_t dbuf = 0;
};

// Group arguments of a launch: gd == nullptr means one group described by
// the launch's own shape arguments.
struct GroupArgs {
  const GroupDev* gd = nullptr;      // device table (G entries)
  const GroupDev* host = nullptr;    // host mirror of the table (launch-time maxima)
  int G = 1;
  unsigned nbmax = 0;                // max STF beams per group (rows of the RMS table)
  size_t prow = 0;                   // baseline points per row (max nbeams * sample_count)
};

// Fills the baseline and per-beam RMS fields of sh (nbeams, stride, nout and
// bc set by the caller); returns an hd_error code.
int series_shape(float cur_dt, float baseline_length, SeriesShape& sh);

// Zero-DM RFI cleaning (clean_filterbank_rfi.cu:256-528): geometry and the
// byte offsets of its device workspace.
struct CleanPlan {
  size_t nsamps, nbits, nchans, stride_w;  // stride_w: 32-bit words per time sample
  float dm, rfi_tol;
  size_t boxcar_max, max_delay, nsmooth;
  SeriesShape sh;                          // the cleaning series (nout = nsamps - max_delay)
  std::vector<int> delays;                 // per-channel delays at dm
  size_t brms_words;
  size_t off_series, off_S, off_pts, off_rms, off_brms, off_mask, off_delays, work_bytes;
};
int clean_plan(size_t nsamps, size_t nbits, size_t nchans, size_t nbeams, float dm, float dt,
               float f0, float df, float baseline_length, float rfi_tol, size_t boxcar_max,
               CleanPlan& cp);
// Writes the first nsamps - max_delay time samples of d_out (d_out != d_in).
hipError_t run_clean(const CleanPlan& cp, const uint8_t* d_in, uint8_t* d_out, void* d_work,
                     hipStream_t s);

// ---- launchers (return hipError_t of the launch) ----
// live_is_identity: live channel l is channel l (no channel masked).
// trim_rows: a group's rows are written only as far as its dedispersion reads
// (nsamps plus two time tiles), not out to xpitch: for ragged batches whose
// rows nothing else reads.
hipError_t launch_unpack(const uint8_t* fb, size_t stride_bytes, int nbits, size_t nsamps,
                         const int* d_live, int n_live, uint8_t* x, size_t xpitch,
                         hipStream_t s, const GroupArgs& ga = GroupArgs{},
                         bool live_is_identity = false, bool trim_rows = false);
// Halves the time resolution of unpacked rows: y[l][t] = (x[l][2t] +
// x[l][2t+1]) >> 1 for t < nsamps / 2, zero from there to byte `limit` of the
// row (0: the pitch; a multiple of 4, at least nsamps / 2); both pitches
// xpitch, a multiple of 256 and at least nsamps + 8.
hipError_t launch_scrunch2(const uint8_t* x, int n_live, size_t xpitch, size_t nsamps,
                           uint8_t* y, hipStream_t s, size_t limit = 0);

struct DedispGeom {
  int n_dm, n_live;
  size_t nout;         // outputs per DM
  size_t xpitch;       // bytes per unpacked channel row
  size_t ppitch;       // floats per output row
  int wlen;            // LDS dwords per staged copy (the largest block's need)
  int db = 16;         // trials per block (8, 16 or 32; the staging pairs are planned for it)
  int remap = 1;       // XCD-aware block order
  int widen8 = 0;      // > 0: byte-staged kernel, widening its lanes every widen8 channels
  int wlen8 = 0;       // its LDS dwords per staged copy
  int split8 = 0;      // byte-staged kernel for 8-bit samples (even/odd byte sums)
  int cc = 4;          // channels staged per step (4, 8 or 16; the staging pairs are planned for it)
  int wide = 0;        // byte staging at 32 trials: 1024 threads (2 trials per wave) instead of 512
  size_t tiles = 0;    // time tiles of the grid (0: ppitch / tile time); must reach every
                       // group's tile holding its sample nout
  // byte staging at 32 trials on 1024 threads: DM blocks 0 .. dbuf_blocks - 1
  // (each block's trials and window of at most wlen8_dbuf dwords the same in
  // every group) run with two staging buffers; 0: none
  int dbuf_blocks = 0;
  int wlen8_dbuf = 0;
  // instead, per group: the groups whose GroupDev::dbuf is set (every DM
  // block's window at most wlen8_dbuf dwords) run with two buffers
  int dbuf_groups = 0;
  int mb_off = 0;      // (set by the launcher: the first DM block of a launch)
};
hipError_t launch_dedisperse(const uint8_t* x, const int* d_delays, const int* d_chunk_pairs,
                             const DedispGeom& g, float* plane, hipStream_t s,
                             const GroupArgs& ga = GroupArgs{});
size_t dedisp_lds_bytes(const DedispGeom& g);  // LDS per block of g's tile (limit 160 KiB)
// the same with two staging buffers of wlen8 dwords per copy
size_t dedisp_lds_bytes_dbuf(const DedispGeom& g, int wlen8);
int dedisp_tile_time();     // outputs per block in time
int dedisp_tile_dms();      // default DMs per block
int dedisp_window_dwords(int spread);  // staged dwords per copy for a delay span
int dedisp8_window_dwords(int spread); // the same for the 8-bit-lane kernel
int dedisp_chunk_channels();           // default channels staged per step
// A byte-staged tile (widen8 > 0, remap 0) whose staged copies fit the LDS for
// the delays d (rows x nl, blocks from row 0) of one launch: 32 trials x 8
// channels on 1024 threads first, then 32 x 4, 16 x 4 and 8 x 4.  False when
// none fits.
bool pick_byte_tile(const std::vector<int>& d, size_t rows, size_t nl, size_t nbits,
                    DedispGeom& dg);

hipError_t launch_baseline_points(const float* plane, const SeriesShape& sh, int n_dm,
                                  float* pts, hipStream_t s, const GroupArgs& ga = GroupArgs{});
// Per-beam RMS of every series; with scratch (beam_rms_scratch_words floats)
// large series take the two-phase kernels.
size_t beam_rms_scratch_words(const SeriesShape& sh, int n_dm,
                              const GroupArgs& ga = GroupArgs{});
hipError_t launch_beam_rms(const float* plane, const float* pts, const SeriesShape& sh,
                           int n_dm, float* rms, hipStream_t s, float* scratch = nullptr,
                           size_t scratch_words = 0, const GroupArgs& ga = GroupArgs{});
hipError_t launch_normalise(float* plane, const float* pts, const float* rms,
                            const SeriesShape& sh, int n_dm, hipStream_t s);
// Prefix sums of n_series rows; with rms (and pts, sh) the rows are first
// normalised as launch_normalise would.
// With ga.gd the rows are the groups' (group, trial) rows, n_series trials each.
hipError_t launch_scan(const float* z, size_t zpitch, size_t count, int n_series, float* S,
                       size_t spitch, hipStream_t s, const float* pts = nullptr,
                       const float* rms = nullptr, const SeriesShape* sh = nullptr,
                       const GroupArgs& ga = GroupArgs{});

// Block RMS, thresholding and giant assembly run over the groups of ga
// (ga.gd required); partials are laid out with np_max = max over groups.
size_t block_rms_np_max(const GroupArgs& ga);
hipError_t launch_block_rms_partials(const float* S, size_t spitch, const WidthTable& wt,
                                     int first_w, int n_w_group, int n_dm, float* partial,
                                     const float* winv, hipStream_t s, const GroupArgs& ga);
// Without ga.gd, `single` gives the one shape (rows 0..n_dm-1).
hipError_t launch_block_rms_finish(const float* partial, int nw_total, int first_w,
                                   int n_w_group, int n_dm, float* wrms, float* winv,
                                   hipStream_t s, const GroupArgs& ga,
                                   const BlockRmsShape* single = nullptr);

struct GiantBuffers {
  uint64_t* skey;      // (dm*NW+w)<<32 | i   (sample records, ordered later)
  float* sval;
  size_t scap;         // sample capacity
  unsigned* counter;   // samples reserved (atomic)
  unsigned* ccount;    // per chunk key
  unsigned* coff;      // per chunk key: offset in the unordered buffer
  unsigned* cpos;      // per chunk key: offset in the ordered buffer
  uint64_t* okey;      // ordered samples
  float* oval;
  unsigned* head;      // per ordered sample
  unsigned* segid;
  unsigned* scan_tmp;
  size_t nchunks_per_w;
  size_t chunk;        // outputs per chunk
};
// prefix-sum positions per launch_threshold block (one key each) for width table wt
size_t threshold_chunk(const WidthTable& wt);
hipError_t launch_threshold(const float* S, size_t spitch, const WidthTable& wt,
                            const float* winv, int n_dm, float thresh, GiantBuffers& gb,
                            hipStream_t s, const GroupArgs& ga);
// p[0..n) = v with a kernel (no memset node when a stream is captured).
hipError_t launch_fill_u32(unsigned* p, size_t n, unsigned v, hipStream_t s);
hipError_t launch_exclusive_scan_u32(const unsigned* in, unsigned* out, size_t n,
                                     unsigned* tmp, hipStream_t s);
size_t exclusive_scan_tmp_words(size_t n);
hipError_t launch_order_samples(GiantBuffers& gb, size_t nkeys, hipStream_t s);
// The samples of launch_threshold (keys of gb.chunk positions) in key order,
// one thread per sample.
hipError_t launch_order_samples(GiantBuffers& gb, const WidthTable& wt, unsigned nsamp,
                                hipStream_t s);
hipError_t launch_segment_heads(const GiantBuffers& gb, const WidthTable& wt, unsigned nsamp,
                                hipStream_t s);
// giants SoA outputs
struct GiantOut {
  float* peak;
  uint64_t* ind;
  uint64_t* begin;
  uint64_t* end;
  uint32_t* filter;
  uint32_t* dm;
  uint32_t* group;     // group of each giant (nullable)
};
// With gd (the launch's group table), a giant's trial is its group's trial0
// plus its row and its group is the group's cgroup.
hipError_t launch_segment_reduce(const GiantBuffers& gb, const WidthTable& wt, unsigned nsamp,
                                 unsigned ngiants, int n_dm, GiantOut go, uint64_t* best,
                                 hipStream_t s, const GroupDev* gd = nullptr);

// clustering / merge
struct ClusterBuffers {
  uint64_t* sort_keys;   // padded power of two
  uint32_t* lo;
  uint32_t* hi;
  int32_t* labels;
  unsigned* flag;
  uint64_t* best;
  unsigned* members;
  unsigned* root_flag;
  unsigned* root_pos;
  unsigned* scan_tmp;
  size_t sort_n;
  // giants' fields in the sorted order (optional): the window scans then read
  // contiguous memory
  uint64_t* s_ind = nullptr;
  uint32_t* s_fd = nullptr;    // dm << 5 | filter
  unsigned band_cap = 0;       // lo/hi entries per giant (filter-banded windows)
  // coincidence lists (optional, banded windows only): for each giant the
  // earlier giants it coincides with, built once so that the relaxation
  // passes read labels only.  ecount and eoff hold n + 1 entries; adj_alloc
  // returns room for the given number of entries (null: windows instead).
  uint32_t* ecount = nullptr;
  uint32_t* eoff = nullptr;
  std::function<uint32_t*(size_t)> adj_alloc;
  // clustering groups (optional, with the lists): n_cgroups groups (the
  // gulps of a call, or 1 without a group column) whose giants are
  // contiguous in index order; gbounds holds 2 * n_cgroups + 3 entries.  The
  // lists are then relaxed per group, one block each with the group's labels
  // in LDS, when every group fits there.
  uint32_t* gbounds = nullptr;
  unsigned n_cgroups = 0;
};
// With group (per giant) and gd, giants of different groups never coincide
// and each group uses its own beam length; otherwise nsamps_beam applies.
// *iterations: relaxation passes run between host checks of the change flag
// (0 when a small search ran its passes inside one block).
hipError_t launch_cluster(const uint64_t* ind, const uint32_t* filter, const uint32_t* dm,
                          const uint64_t* members_in, unsigned n, uint64_t time_tol,
                          uint64_t filter_tol, uint64_t dm_tol, uint64_t nsamps_beam,
                          unsigned max_filter, ClusterBuffers& cb, hipStream_t s,
                          unsigned* h_flag_pinned, int* iterations,
                          const uint32_t* group = nullptr, const GroupDev* gd = nullptr,
                          uint64_t max_coord = ~0ull);
hipError_t launch_merge(const float* peak, const uint64_t* ind, const uint64_t* begin,
                        const uint64_t* end, const uint32_t* filter, const uint32_t* dm,
                        const uint64_t* members_in, unsigned n, ClusterBuffers& cb,
                        float* g_peak, uint64_t* g_ind, uint64_t* g_begin, uint64_t* g_end,
                        uint64_t* g_filter, uint64_t* g_dm, uint64_t* g_members,
                        hipStream_t s, const uint32_t* group = nullptr,
                        uint32_t* g_group = nullptr);

// ---- learned proposal (k_proposal.hip, k_cnn.hip) ----
// Coarse input: x (slot g of each group, xpitch) summed over 16 samples into
// xc (xpitch_c): (sum + 8) >> 4 for 8-bit samples, the plain sum otherwise.
// gd[g].nsamps is the group's coarse sample count.
hipError_t launch_decimate(const uint8_t* x, size_t xpitch, int n_live, int nbits, uint8_t* xc,
                           size_t xpitch_c, const GroupDev* gd, int G, hipStream_t s);
// Normalised coarse plane as bfloat16, bf[(g*n_rows + m)*bpitch + t], and
// its wide plane bw[(g*n_rows + m)*wpitch + j] = (z[4j] + .. + z[4j+3]) / 2.
hipError_t launch_coarse_bf16(const float* plane, size_t ppitch, const float* pts, size_t prow,
                              const float* rms, unsigned nbmax, const GroupDev* gd, int G,
                              int n_rows, uint16_t* bf, size_t bpitch, uint16_t* bw,
                              size_t wpitch, hipStream_t s);
// Two-channel tiles [64][64][2] of the coarse and wide planes.  Explicit: n
// tiles at (u, v) = pos[2i..2i+1] of
// group 0.  Enumerated: tiles first .. first+n-1 of the order (g, u, v),
// tile_off[g] the first tile of group g (G+1 entries), n_tpos[g] its time
// positions, n_units units per group.
struct TileGrid {
  const uint16_t* bf;
  size_t bpitch;
  const uint16_t* bw;         // wide plane
  size_t wpitch;
  int n_rows;                 // coarse rows per group
  const uint32_t* ncols;      // coarse samples of each group's plane
  int n_units;
  const uint32_t* tile_off;   // enumerated mode
  const uint32_t* n_tpos;
  int G;
};
hipError_t launch_extract_tiles(const TileGrid& tg, const uint32_t* pos, size_t first, size_t n,
                                uint16_t* tiles, hipStream_t s);
// flags[g*n_units + u] = 1 if any tile (g, u, v) has score > thresh.
hipError_t launch_unit_flags(const float* scores, const uint32_t* tile_off,
                             const uint32_t* n_tpos, int G, int n_units, float thresh,
                             uint8_t* flags, hipStream_t s);

}  // namespace hd

// The proposal network (k_cnn.hip): forward of n tiles on stream s.
struct hd_cnn_t;
namespace hd {
hipError_t cnn_forward_async(hd_cnn_t* net, const uint16_t* tiles, size_t n, float* logits,
                             hipStream_t s);
size_t cnn_max_batch(hd_cnn_t* net);
}  // namespace hd
