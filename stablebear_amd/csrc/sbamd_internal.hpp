// Internal types shared by the host runtime and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/stablebear_amd.h"

namespace sbamd
{
  // ---------------------------------------------------------------------
  // Error plumbing: C++ exceptions never cross the C-ABI; api.cpp converts
  // them into status codes + a thread-local message.
  // ---------------------------------------------------------------------
  struct Error
  {
    int code;
    std::string msg;
  };

  [[noreturn]] void fail(int code, const std::string& msg);
  void hip_check(hipError_t e, const char* what, const char* file, int line);

#define SBAMD_HIP(call) ::sbamd::hip_check((call), #call, __FILE__, __LINE__)

  // ---------------------------------------------------------------------
  // Device layout
  // ---------------------------------------------------------------------

  // One workgroup tile: rows [r0, r0 + nr) x cols [c0, c0 + nc) in the sorted
  // PCF order of problem `prob`.
  struct Tile
  {
    int32_t r0;
    int32_t c0;
    int16_t nr;
    int16_t nc;
    uint16_t prob;
    uint16_t flags; // bit 0: diagonal tile (rows == cols); bit 1: fold partials in FP64
  };
  static_assert(sizeof(Tile) == 16, "Tile must stay 16 bytes");

  // Tile kernel geometry. With `ppl` pairs per lane, lane (r, c) of a tile
  // walks pairs (r, c + q * ceil(C / ppl)), q < ppl; lanes past the tile's
  // pairs walk a stand-in pair and store nothing. Workgroups are whole
  // wavefronts of at most 512 lanes; one wavefront reduces the R + C slots'
  // phase-end proposals, so R + C <= 64.
  __host__ __device__ inline int tile_lane_cols(int C, int ppl) { return (C + ppl - 1) / ppl; }
  __host__ __device__ inline int tile_block_threads(int R, int C, int ppl)
  {
    const int lanes = R * tile_lane_cols(C, ppl);
    return lanes <= 64 ? 64 : (lanes + 63) / 64 * 64;
  }
  inline bool tile_shape_ok(int R, int C, int cap, int ppl)
  {
    return (ppl == 1 || ppl == 2) && R >= 1 && C >= 1 && R + C <= 64 &&
           tile_block_threads(R, C, ppl) <= 512 && cap >= 2 && cap <= 4096;
  }

  constexpr uint16_t TILE_DIAG = 1;
  constexpr uint16_t TILE_FOLD = 2;

  // Kernel-level pair operations (SBAMD_OP_LP with p == 2 becomes OPK_L2).
  constexpr int OPK_L1 = 0;
  constexpr int OPK_LP = 1;
  constexpr int OPK_DOT = 2;
  constexpr int OPK_L2 = 3;

  // Per-problem bases into the concatenated arrays of a (batched) set.
  struct Problem
  {
    int64_t row_base; // first row PCF (sorted order) of the problem in r_off / r_perm
    int64_t col_base; // first column PCF of the problem in c_off / c_perm
    int64_t out_base; // first output element of the problem
    int64_t n;        // PCFs in the problem (rows)
  };

  // Kernel parameters of the forward lane-per-pair tile kernel.
  // Work queues of a persistent launch, one per XCD.
  constexpr int DEVICE_QUEUES = 8;

  struct FwdParams
  {
    const int64_t* r_off;  // record offsets of the row set (sorted PCF order), n+1 per problem base
    const void* r_rec;     // records {t_end (TT), value (TV)}, 16 bytes (FP64 times) or 8
    const int64_t* r_perm; // sorted position -> original index within its problem
    const int64_t* c_off;
    const void* c_rec;
    const int64_t* c_perm;
    const Tile* tiles;
    const Problem* probs;
    void* out;
    double* row_sum; // optional (DENSE): per original row sums
    int64_t ntiles;
    int64_t out_cols; // DENSE row length
    double p;
    int32_t out_kind;
    int32_t out_f64;
    int32_t R, C, cap, stride;
    int32_t ppl; // pairs per lane (1 or 2)
    int32_t symmetric; // rows and cols are the same set (pdist / kernel)
    int32_t include_diag;
    uint32_t* work_counter; // non-null: persistent workgroups take tiles from these counters (64 bytes apart)
    // persistent: queue x holds tiles q_tab[x] .. q_tab[x] + q_tab[8 + x] - 1;
    // a workgroup takes from the queue of its XCD, then from the others (null:
    // every tile in queue 0)
    const int64_t* q_tab;
  };

  // Kernel parameters of the long-pair kernel (one wavefront per pair).
  struct LongParams
  {
    const int64_t* r_off;
    const void* r_rec;
    const int64_t* r_perm;
    const int64_t* c_off;
    const void* c_rec;
    const int64_t* c_perm;
    void* out;
    double* row_sum;  // optional (DENSE): per original row sums
    int64_t q0, q1;   // pair range [q0, q1) in the plan's pair order
    int64_t n_cols;   // DENSE row length (cdist)
    double p;
    int32_t out_kind;
    int32_t out_f64;
  };

  // Kernel parameters of the per-PCF norm kernel.
  struct NormParams
  {
    const int64_t* roff; // sorted record offsets (n + 1)
    const void* rec;     // records
    const int64_t* perm; // sorted position -> original index
    int64_t n;
    double p;
    void* out; // n norms in original order
    int32_t out_f64;
    int32_t pad;
  };

  // One event of a group's merged breakpoint list (a group is 64 consecutive
  // PCFs): the record at position pos starts at time t, dt after the group's
  // previous event, and the PCF's value jumps there by d. col is the PCF's
  // index within the group, with BWD_FIRST set on a PCF's first record.
  struct BwdEvent
  {
    double t;   // record start time
    double d;   // value jump at t (the first record: its value)
    double dt;  // d * t
    int32_t col; // PCF within its group (| BWD_FIRST for a first record)
    int32_t pos; // record position
  };
  constexpr int32_t BWD_GROUP = 64;
  constexpr int32_t BWD_FIRST = 0x40000000;
  constexpr int64_t BWD_EVENT_PAD = 8; // events readable past the end of the sublist array
  static_assert(sizeof(BwdEvent) == 32, "BwdEvent must stay 32 bytes");

  // Parameters of the backward kernels (kernels_bwd.hip). Arrays marked
  // "position" are in the original PCF order: record k of PCF i at ooff[i] + k.
  struct BwdParams
  {
    int64_t n;       // PCFs
    int64_t records; // records of the set (= positions)
    int32_t ch;      // partners per chunk: 64 (FP32 pair weights) or 32 (FP64)
    int32_t op;      // OPK_L2 (pdist) or OPK_DOT (kernel with diagonal)
    int32_t in_f64;  // grad_out / fwd_out are FP64
    int32_t w_f64;   // pair weights kept in FP64 (else FP32) in LDS
    int32_t val_f64; // record values are FP64 (else FP32); record times are FP64
    int32_t grad_f64;
    // the sorted device set
    const int64_t* roff;
    const void* rec;
    const int64_t* perm;
    const int64_t* src;
    // position layout
    const int64_t* ooff; // n + 1
    double* key0;        // record start time
    double* val0;        // record value
    int32_t* col0;       // owning PCF
    int64_t* src0;       // source point (-1: zero record of an empty PCF)
    double* keys[2];     // merge sort ping-pong
    int32_t* ids[2];
    BwdEvent* events;  // event lists of the groups of BWD_GROUP PCFs (the bands)
    BwdEvent* events2; // event lists of the groups of ch / 2 PCFs (the chunks' sublists), with
                       // BWD_EVENT_PAD events of padding after the last
    double* pst0;      // per position: F(t_start) - v t_start (set inside launch_backward)
    // loss and accumulators
    const void* grad_out;
    const void* fwd_out;
    double* q;    // per position: S = integral of H from 0 to the record's start
    double* a;    // per position: H where the record ends (last record: S at DBL_MAX)
    double* wsum; // per PCF: sum of its pair weights
    void* grad_v;
    void* grad_t;
    // set by the list build: 1 if two different PCFs share a breakpoint time
    // (> 0), else 0; NULL: not known, every walk checks for ties
    int32_t* cross_ties;
  };

  // ---------------------------------------------------------------------
  // Host objects behind the opaque C handles
  // ---------------------------------------------------------------------
  struct DeviceBuffer
  {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t b);
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept;
    ~DeviceBuffer();
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
  };

  // Host-side packed set (validated, zero-length segments dropped, sorted).
  struct PackedSet
  {
    int64_t n = 0;          // PCFs
    int64_t points = 0;     // original points (offsets[n])
    int32_t in_dtype = SBAMD_F64;
    std::vector<int64_t> perm;   // sorted position -> original PCF index
    std::vector<int64_t> roff;   // sorted record offsets (n + 1)
    std::vector<int64_t> src;    // record -> original flat point index
    std::vector<double> tend64;  // record end times (inf for the last record)
    std::vector<double> val64;   // record values
  };

  void pack_set(const sbamd_pcf_set& in, PackedSet& out);
  void validate_set(const sbamd_pcf_set& in);
  void check_set_offsets(int64_t n, const int64_t* offsets, int32_t dtype);
  std::vector<int64_t> record_counts(const sbamd_pcf_set& in);
  void sort_by_records(const std::vector<int64_t>& m, std::vector<int64_t>& perm);

  // Pinned host buffers that host-to-host jobs stage their transfers
  // through, kept for the next job: pinning (and first touching) hundreds of
  // MB costs more than the copies they serve. At most KEEP idle buffers are
  // held; the smallest goes first.
  struct PinnedPool
  {
    static constexpr size_t KEEP = 4;
    std::mutex mu;
    std::vector<std::pair<void*, size_t>> idle;
    PinnedPool() = default;
    PinnedPool(const PinnedPool&) = delete;
    PinnedPool& operator=(const PinnedPool&) = delete;
    ~PinnedPool();
    void* take(size_t bytes, size_t& cap); // an idle buffer of at least `bytes`, else a new one
    void give(void* p, size_t cap);
  };

  // One pinned buffer of a pool for the lifetime of the lease.
  struct PinnedLease
  {
    PinnedPool* pool = nullptr;
    void* ptr = nullptr;
    size_t cap = 0;
    PinnedLease() = default;
    PinnedLease(PinnedPool& pl, size_t bytes) : pool(&pl) { ptr = pl.take(bytes, cap); }
    PinnedLease(const PinnedLease&) = delete;
    PinnedLease& operator=(const PinnedLease&) = delete;
    ~PinnedLease()
    {
      if (ptr)
        pool->give(ptr, cap);
    }
  };

  struct Ctx
  {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    size_t lds_per_block = 0;
    void* scratch = nullptr; // this device's address of the no-alloc working memory
    std::string arch;        // gcnArchName (autotuner cache key)
    // host-to-host jobs: output bytes the device may hold at once (0: from
    // free memory) and the tile shape of their plans (0 fields: defaults)
    uint64_t out_budget = 0;
    int32_t job_R = 0, job_C = 0, job_cap = 0, job_ppl = 0;
    PinnedPool pinned;
  };

  struct Set
  {
    Ctx* ctx = nullptr;
    int32_t precision = SBAMD_PREC_FP64;
    int32_t in_dtype = SBAMD_F64;
    int32_t tt_f64 = 1; // record times stored as FP64
    int32_t tv_f64 = 1; // record values stored as FP64
    int64_t n = 0;
    int64_t records = 0;
    int64_t points = 0;
    std::vector<int64_t> h_roff; // sorted record offsets
    std::vector<int64_t> h_perm;
    bool has_empty = false; // some PCF has no points (its record has no source point)
    uint64_t vgen = 0;      // bumped by every value update (batched plans copy records)
    // rec: one record per constant piece, {t_end, value} in the set's time and
    // value types (16 bytes with FP64 times, 8 with FP32 times), sorted order
    DeviceBuffer roff, rec, perm, src;
  };

  // Length classes of the per-class tile shapes: a PCF's class is set by
  // its record count (thresholds 128, 512, 2048). Sets are sorted by record
  // count, so a class is a range of sorted positions.
  constexpr int NCLASS = 4;
  inline int length_class(int64_t m) { return m < 128 ? 0 : (m < 512 ? 1 : (m < 2048 ? 2 : 3)); }
  constexpr int64_t CLASS_MIN[NCLASS + 1] = {0, 128, 512, 2048, INT64_MAX};

  struct Shape
  {
    int32_t R, C, cap, ppl;
  };

  struct Plan
  {
    Ctx* ctx = nullptr;
    std::vector<const Set*> sets; // 1 (pdist / kernel), 2 (cdist: rows, cols), or k (batched)
    bool cdist = false;
    bool batched = false;
    int32_t op = SBAMD_OP_L1;
    double p = 1.0;
    int32_t out_kind = SBAMD_OUT_DISTMAT;
    int32_t precision = SBAMD_PREC_FP64;
    int32_t tt_f64 = 1, tv_f64 = 1;
    // tile shape (set_default_shape gives a new plan's; the autotuner may
    // pick another)
    int32_t R = 16, C = 16, cap = 48;
    int32_t ppl = 1; // pairs per lane of the tile kernel
    // per length class shapes (set by the autotuner): the tiles of class k
    // are h_tiles[class_begin[k], class_begin[k + 1]) and run in a launch of
    // their own; without them every tile has the shape above, in one launch
    bool per_class = false;
    Shape cls[NCLASS] = {};
    int64_t class_begin[NCLASS + 1] = {};
    // persistent launches: class k's tiles in per-XCD queues, queue x at
    // class_begin[k] + xq[k][x] .. class_begin[k] + xq[k][x + 1] (all in
    // queue 0 when the tiles were not ordered so)
    int64_t xq[NCLASS][DEVICE_QUEUES + 1] = {};
    DeviceBuffer qtab; // per class: the queues' starts, then their lengths (FwdParams::q_tab)
    Shape shape_of(int k) const { return per_class ? cls[k] : Shape{R, C, cap, ppl}; }
    int64_t fold_threshold = 8192;
    uint64_t pairs = 0, segments = 0, out_count = 0;
    std::vector<Problem> h_probs;
    std::vector<Tile> h_tiles;
    DeviceBuffer probs, tiles;
    // batched: concatenated set arrays; member b's records start at record
    // b_at[b] of b_rec, copied when the member's values were at generation
    // b_vgen[b]
    DeviceBuffer b_roff, b_rec, b_perm;
    std::vector<int64_t> b_at;
    std::vector<uint64_t> b_vgen;
    // persistent launch: resident workgroups pull tiles from this counter
    bool persistent = false;
    DeviceBuffer counter;
    // backward working arrays (kernels_bwd.hip), allocated on first use
    DeviceBuffer bwd_ooff, bwd_key0, bwd_val0, bwd_col0, bwd_src0, bwd_keys[2], bwd_ids[2],
        bwd_events, bwd_events2, bwd_q, bwd_a, bwd_w, bwd_ties;
    // the event lists in bwd_events / bwd_events2 are sorted (a set's
    // breakpoint times never change, so they stay sorted until reallocated)
    bool bwd_lists_ready = false;
    // long-pair kernel: mode 0 off, 1 on, 2 auto (few long pairs, FP32C);
    // use_long is the decision for this plan
    int32_t long_mode = 2;
    bool use_long = false;
    uint64_t input_bytes = 0;
    // Launch state generation: changes whenever the device buffers or launch
    // parameters a captured graph holds are replaced (a unique number per
    // plan state across the process).
    uint64_t gen = 0;
    void bump();
    void rebuild_tiles();
  };

  // Kernel launchers (kernels_fwd.hip / kernels_bwd.hip / kernels_misc.hip)
  size_t fwd_lds_bytes(int tt_f64, int tv_f64, int R, int C, int cap);
  void launch_forward(const FwdParams& prm, int tt_f64, int tv_f64, int op, int kahan,
                      hipStream_t stream);
  void launch_long(const LongParams& prm, int tt_f64, int tv_f64, int op, int kahan, int cu_count,
                   hipStream_t stream);
  void launch_norm(const NormParams& prm, int tt_f64, int tv_f64, hipStream_t stream);
  // lists_ready: the event lists of this set were sorted by an earlier
  // backward on the same working arrays (only their value fields are redone)
  void launch_backward(const BwdParams& prm, bool lists_ready, hipStream_t stream);
  void launch_fill_zero(void* ptr, size_t bytes, hipStream_t stream);
  void launch_gather_subset(const int64_t* roff, const void* rec, int rec_bytes, const int64_t* sel,
                            const int64_t* droff, int64_t n, int64_t records, void* drec, hipStream_t stream);
  constexpr int FILL_MAX = 8;
  void launch_fill_zero_multi(void* const* ptrs, const size_t* bytes, int n, hipStream_t stream);

  // Device-side packing of a strided CSR view (kernels_pack.hip).
  struct PackParams
  {
    int64_t n;                 // PCFs
    const int64_t* off;        // n + 1 point offsets (device)
    const unsigned char* t;    // point k's time at t + k * t_stride
    const unsigned char* v;    // point k's value at v + k * v_stride
    int64_t t_stride, v_stride;
    int64_t* m;                // pass 1: records per PCF (original order)
    unsigned long long* bad;   // pass 1: lowest PCF index breaking a rule
    const int64_t* roff;       // pass 2: sorted record offsets (n + 1)
    const int64_t* qinv;       // pass 2: PCF -> sorted position
    const int64_t* pbase;      // optional: PCF i's points are at byte pbase[i] + k * stride
    void* rec;                 // pass 2: records
    int64_t* src;              // pass 2: record -> point (-1: zero PCF)
    int32_t tt_f64, tv_f64;    // record types
  };
  void launch_pack_count(const PackParams& P, int f64, hipStream_t stream);

  // Row reductions of a pairwise output in device memory (kernels_rows.hip).
  struct RowsParams
  {
    const void* d;        // dense n_rows x n_cols, or compressed lower triangle (n_rows = n)
    int64_t n_rows, n_cols;
    int32_t compressed;   // 1: distance-matrix storage, rows over j != i
    int32_t f64;          // element type
    double* sum;          // optional per-row outputs (FP64)
    double* min;
    double* max;
    int32_t k;            // k nearest: values (element type) and columns
    void* knn_v;
    int64_t* knn_i;
  };
  void launch_rows_reduce(const RowsParams& P, hipStream_t stream);
  void launch_rows_knn(const RowsParams& P, hipStream_t stream);

  // The reference's file format (io.cpp).
  void io_parse(const void* buf, uint64_t len, sbamd_io_info& info);
  uint64_t io_index(const void* buf, uint64_t len, const sbamd_io_info& info, uint64_t* counts,
                    uint64_t* starts);
  uint64_t io_header(uint32_t kind, int32_t base_format, int32_t sub_format, int32_t ndim,
                     const uint64_t* shape, void* out, uint64_t cap);
  uint64_t io_write_pcfs(int64_t n, const int64_t* offsets, const void* t, const void* v, int32_t dtype,
                         void* out, uint64_t cap);
  void launch_pack_records(const PackParams& P, int f64, hipStream_t stream);
  void launch_rectangles(const double* ft, const double* fv, int64_t mf, const double* gt,
                         const double* gv, int64_t mg, double a, double b, double* out,
                         int64_t cap, int64_t* count, hipStream_t stream);
  void launch_rectangles_f32(const float* ft, const float* fv, int64_t mf, const float* gt,
                             const float* gv, int64_t mg, float a, float b, float* out,
                             int64_t cap, int64_t* count, hipStream_t stream);
  // Writes new values into the value field of a record array (records of
  // rec_bytes bytes, the value at byte offset rec_bytes / 2).
  void launch_update_values(const void* v, const int64_t* src, int64_t records, void* rec,
                            int rec_bytes, int in_f64, int out_f64, hipStream_t stream);

  // Bytes of one device record for the given time type.
  inline int record_bytes(int tt_f64) { return tt_f64 ? 16 : 8; }
  void launch_plain(const int64_t* off, const double* t, const double* v, int64_t n, int op,
                    double p, int out_kind, double* out, hipStream_t stream);

  // Planning
  void build_tiles_region(const std::vector<int64_t>& m_rows, const std::vector<int64_t>& m_cols,
                          int64_t r_lo, int64_t r_hi, int64_t c_lo, int64_t c_hi, int R, int C, bool tri,
                          bool include_diag, uint16_t prob, int64_t fold_threshold, std::vector<Tile>& tiles);
  void build_tiles(const std::vector<int64_t>& m_rows_sorted, const std::vector<int64_t>* m_cols_sorted,
                   int R, int C, bool include_diag, uint16_t prob, int64_t fold_threshold,
                   std::vector<Tile>& tiles);
  void sort_by_records(const std::vector<int64_t>& m, std::vector<int64_t>& perm);
  int xcd_supertile_order(std::vector<Tile>& tiles, size_t b0, size_t b1, int R, int C, int nxcd);
  // Per-XCD work queues of a persistent launch over tiles [b0, b1) (given in
  // longest-first order, `work` per tile): units of G x G tiles of one
  // problem, which share their PCFs, are dealt largest first to the queue
  // with the least work so far; the queues are laid out one after the other,
  // queue x at [b0 + qoff[x], b0 + qoff[x + 1]).
  void xcd_queue_order(std::vector<Tile>& tiles, size_t b0, size_t b1, int R, int C, int nxcd,
                       const std::vector<int64_t>& work, int64_t* qoff);

  // XCDs of the device (gfx950: 8, each with its own 4 MB L2).
  constexpr int DEVICE_XCDS = 8;
}
