// C-ABI implementation: contexts, device-resident sets, plans, asynchronous
// jobs, graphs, autotuning and the single-pair helpers. No exception crosses
// the boundary: every entry point converts failures into a status code and a
// thread-local message.
#include "sbamd_internal.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <tuple>
#include <type_traits>
#include <unordered_map>

// Working memory of sbamd_plan_backward_noalloc (layout in NOALLOC_* below).
static __device__ __align__(256) unsigned char scratch[(size_t(1) << 19) * 132 + 8 * 32 + (size_t(1) << 16) * 8];

namespace sbamd
{
  [[noreturn]] void fail(int code, const std::string& msg) { throw Error{code, msg}; }

  void hip_check(hipError_t e, const char* what, const char* file, int line)
  {
    if (e == hipSuccess)
      return;
    std::string msg = std::string(file) + ":" + std::to_string(line) + ": " + what + ": " +
                      hipGetErrorString(e);
    fail(e == hipErrorOutOfMemory ? SBAMD_ERR_NOMEM : SBAMD_ERR_RUNTIME, msg);
  }

  DeviceBuffer::DeviceBuffer(size_t b) : bytes(b)
  {
    if (b > 0)
      SBAMD_HIP(hipMalloc(&ptr, b));
  }

  DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept
  {
    if (this != &o)
    {
      if (ptr)
        (void)hipFree(ptr);
      ptr = o.ptr;
      bytes = o.bytes;
      o.ptr = nullptr;
      o.bytes = 0;
    }
    return *this;
  }

  DeviceBuffer::~DeviceBuffer()
  {
    if (ptr)
      (void)hipFree(ptr);
  }

  PinnedPool::~PinnedPool()
  {
    for (auto& e : idle)
      (void)hipHostFree(e.first);
  }

  void* PinnedPool::take(size_t bytes, size_t& cap)
  {
    {
      std::lock_guard<std::mutex> lk(mu);
      size_t best = idle.size();
      for (size_t k = 0; k < idle.size(); ++k)
        if (idle[k].second >= bytes && (best == idle.size() || idle[k].second < idle[best].second))
          best = k;
      if (best < idle.size())
      {
        void* p = idle[best].first;
        cap = idle[best].second;
        idle.erase(idle.begin() + static_cast<std::ptrdiff_t>(best));
        return p;
      }
    }
    void* p = nullptr;
    SBAMD_HIP(hipHostMalloc(&p, std::max<size_t>(bytes, 1), hipHostMallocDefault));
    cap = std::max<size_t>(bytes, 1);
    return p;
  }

  void PinnedPool::give(void* p, size_t cap)
  {
    void* drop = nullptr;
    {
      std::lock_guard<std::mutex> lk(mu);
      idle.emplace_back(p, cap);
      if (idle.size() > KEEP)
      {
        auto it = std::min_element(idle.begin(), idle.end(),
                                   [](const auto& a, const auto& b) { return a.second < b.second; });
        drop = it->first;
        idle.erase(it);
      }
    }
    if (drop)
      (void)hipHostFree(drop);
  }

  namespace
  {
    thread_local std::string g_last_error;

    template <typename F>
    int guarded(F&& f)
    {
      try
      {
        f();
        return SBAMD_OK;
      }
      catch (const Error& e)
      {
        g_last_error = e.msg;
        return e.code;
      }
      catch (const std::bad_alloc&)
      {
        g_last_error = "host allocation failed";
        return SBAMD_ERR_NOMEM;
      }
      catch (const std::exception& e)
      {
        g_last_error = e.what();
        return SBAMD_ERR_RUNTIME;
      }
      catch (...)
      {
        g_last_error = "unknown error";
        return SBAMD_ERR_RUNTIME;
      }
    }

    void need(const void* p, const char* what)
    {
      if (!p)
        fail(SBAMD_ERR_INVALID, std::string(what) + " is NULL");
    }

    void use_device(const Ctx* ctx) { SBAMD_HIP(hipSetDevice(ctx->device)); }

    // Stream arguments are hipStream_t values as given: NULL is the null
    // (legacy default) stream, as everywhere in HIP. The context's own stream
    // is available through sbamd_ctx_stream().
    hipStream_t pick_stream(const Ctx*, void* stream) { return static_cast<hipStream_t>(stream); }

    // Working memory of sbamd_plan_backward_noalloc: one static device array
    // (every GPU the library runs on holds its own copy until the process
    // ends), laid out for a set of at most NOALLOC_MAX_PCFS PCFs and
    // NOALLOC_MAX_RECORDS records: per record start time, value, source
    // point, two merge-sort keys, Q, A (8 bytes each), two 32-byte events,
    // owner and two merge-sort ids (4 bytes each); per PCF its weight sum;
    // BWD_EVENT_PAD events of padding after the second event array.
    constexpr int64_t NOALLOC_MAX_RECORDS = int64_t(1) << 19;
    constexpr int64_t NOALLOC_MAX_PCFS = int64_t(1) << 16;
    constexpr size_t NOALLOC_KEY0 = 0;
    constexpr size_t NOALLOC_VAL0 = NOALLOC_KEY0 + 8 * NOALLOC_MAX_RECORDS;
    constexpr size_t NOALLOC_SRC0 = NOALLOC_VAL0 + 8 * NOALLOC_MAX_RECORDS;
    constexpr size_t NOALLOC_KEYS0 = NOALLOC_SRC0 + 8 * NOALLOC_MAX_RECORDS;
    constexpr size_t NOALLOC_KEYS1 = NOALLOC_KEYS0 + 8 * NOALLOC_MAX_RECORDS;
    constexpr size_t NOALLOC_Q = NOALLOC_KEYS1 + 8 * NOALLOC_MAX_RECORDS;
    constexpr size_t NOALLOC_A = NOALLOC_Q + 8 * NOALLOC_MAX_RECORDS;
    constexpr size_t NOALLOC_EVENTS = NOALLOC_A + 8 * NOALLOC_MAX_RECORDS;
    constexpr size_t NOALLOC_EVENTS2 = NOALLOC_EVENTS + sizeof(BwdEvent) * NOALLOC_MAX_RECORDS;
    constexpr size_t NOALLOC_COL0 = NOALLOC_EVENTS2 + sizeof(BwdEvent) * (NOALLOC_MAX_RECORDS + BWD_EVENT_PAD);
    constexpr size_t NOALLOC_IDS0 = NOALLOC_COL0 + 4 * NOALLOC_MAX_RECORDS;
    constexpr size_t NOALLOC_IDS1 = NOALLOC_IDS0 + 4 * NOALLOC_MAX_RECORDS;
    constexpr size_t NOALLOC_W = NOALLOC_IDS1 + 4 * NOALLOC_MAX_RECORDS;
    constexpr size_t NOALLOC_BYTES = NOALLOC_W + 8 * NOALLOC_MAX_PCFS;
    static_assert(sizeof(BwdEvent) == 32, "event layout");
    static_assert(NOALLOC_BYTES == sizeof(scratch), "working memory layout");

    // Host code never passes the array's name to the device: it is a
    // host-side stand-in there. Each device's address of it comes from
    // hipGetSymbolAddress, once per device, with that device current.
    std::mutex g_scratch_mu;
    std::unordered_map<int, void*>& scratch_addr()
    {
      static std::unordered_map<int, void*> m;
      return m;
    }

    // Live plans and their launch-state generations, so that a captured
    // graph can tell whether its plan still holds the buffers and
    // parameters it was captured with.
    std::mutex g_plans_mu;
    std::unordered_map<const Plan*, uint64_t>& live_plans()
    {
      static std::unordered_map<const Plan*, uint64_t> m;
      return m;
    }
    std::atomic<uint64_t> g_plan_gen{0};

    void forget_plan(const Plan* pl)
    {
      std::lock_guard<std::mutex> lk(g_plans_mu);
      live_plans().erase(pl);
    }

    bool plan_current(const Plan* pl, uint64_t gen)
    {
      std::lock_guard<std::mutex> lk(g_plans_mu);
      auto it = live_plans().find(pl);
      return it != live_plans().end() && it->second == gen;
    }

    // Uploads that do not block the host while the device is busy: a
    // blocking hipMemcpy waits for the kernels already queued on the device
    // with some HIP runtimes (the one PyTorch bundles does), which would
    // serialise a job's planning of band k + 1 with band k's kernels. While
    // an AsyncUploads is installed on a thread, that thread's uploads are
    // copied into pinned staging leased from the context's pool and sent
    // with hipMemcpyAsync on the given stream, so they also stay ordered
    // before the kernels later queued on it. The staging is returned when
    // the AsyncUploads ends, which must be after that stream has drained.
    struct AsyncUploads;
    thread_local AsyncUploads* t_async_uploads = nullptr;

    struct AsyncUploads
    {
      static constexpr size_t ARENA = size_t(16) << 20;
      Ctx* ctx;
      hipStream_t stream;
      AsyncUploads* outer;
      std::vector<std::unique_ptr<PinnedLease>> staging;
      size_t used = 0;
      AsyncUploads(Ctx* c, hipStream_t st) : ctx(c), stream(st), outer(t_async_uploads) { t_async_uploads = this; }
      AsyncUploads(const AsyncUploads&) = delete;
      AsyncUploads& operator=(const AsyncUploads&) = delete;
      ~AsyncUploads() { t_async_uploads = outer; }
      void copy(void* dst, const void* src, size_t bytes)
      {
        if (staging.empty() || used + bytes > staging.back()->cap)
        {
          staging.push_back(std::make_unique<PinnedLease>(ctx->pinned, std::max(ARENA, bytes)));
          used = 0;
        }
        unsigned char* at = static_cast<unsigned char*>(staging.back()->ptr) + used;
        std::memcpy(at, src, bytes);
        SBAMD_HIP(hipMemcpyAsync(dst, at, bytes, hipMemcpyHostToDevice, stream));
        used += (bytes + 255) & ~size_t(255);
      }
    };

    template <typename T>
    void upload(DeviceBuffer& buf, const std::vector<T>& v)
    {
      buf = DeviceBuffer(v.size() * sizeof(T));
      if (v.empty())
        return;
      if (t_async_uploads)
        t_async_uploads->copy(buf.ptr, v.data(), v.size() * sizeof(T));
      else
        SBAMD_HIP(hipMemcpy(buf.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    }

    // Kernel op for (op, p): p == 2 takes the d*d / sqrt specialisation.
    int kernel_op(int32_t op, double p)
    {
      if (op == SBAMD_OP_L1)
        return OPK_L1;
      if (op == SBAMD_OP_DOT)
        return OPK_DOT;
      if (op == SBAMD_OP_LP)
        return p == 2.0 ? OPK_L2 : OPK_LP;
      fail(SBAMD_ERR_INVALID, "unknown op");
    }
  }
}

using namespace sbamd;

struct sbamd_ctx : Ctx
{
};
struct sbamd_set : Set
{
};
struct sbamd_plan : Plan
{
};

namespace sbamd
{
  namespace
  {
    void check_shape(int R, int C, int cap, int ppl)
    {
      if (!tile_shape_ok(R, C, cap, ppl))
        fail(SBAMD_ERR_INVALID, "tile shape out of range (rows + cols <= 64, lane_pairs 1 or 2, "
                                "at most 512 lanes, 2 <= cap <= 4096)");
    }

    void set_layout_for(int32_t precision, int32_t in_dtype, int32_t& tt_f64, int32_t& tv_f64)
    {
      if (precision == SBAMD_PREC_FP64)
      {
        tt_f64 = 1;
        tv_f64 = 1;
      }
      else if (precision == SBAMD_PREC_FP32C)
      {
        tt_f64 = in_dtype == SBAMD_F64 ? 1 : 0;
        tv_f64 = 0;
      }
      else
        fail(SBAMD_ERR_INVALID, "unknown precision");
    }

    // Device record array: {t_end, value} per record in the set's types; the
    // value sits at byte offset record_bytes / 2 (FP32 values pad to 16 bytes
    // behind FP64 times, so one record is one aligned 16-byte load).
    std::vector<unsigned char> interleave_records(const PackedSet& pk, int32_t tt_f64, int32_t tv_f64)
    {
      if (!tt_f64 && tv_f64)
        fail(SBAMD_ERR_RUNTIME, "FP64 values need FP64 times");
      const size_t R = pk.tend64.size();
      const size_t rb = static_cast<size_t>(record_bytes(tt_f64));
      std::vector<unsigned char> out(R * rb, 0);
      for (size_t r = 0; r < R; ++r)
      {
        unsigned char* p = out.data() + r * rb;
        if (tt_f64)
          std::memcpy(p, &pk.tend64[r], 8);
        else
        {
          const float t = static_cast<float>(pk.tend64[r]);
          std::memcpy(p, &t, 4);
        }
        if (tv_f64)
          std::memcpy(p + rb / 2, &pk.val64[r], 8);
        else
        {
          const float v = static_cast<float>(pk.val64[r]);
          std::memcpy(p + rb / 2, &v, 4);
        }
      }
      return out;
    }

    void make_set(Ctx* ctx, const sbamd_pcf_set& in, int32_t precision, Set& s)
    {
      PackedSet pk;
      pack_set(in, pk);
      s.ctx = ctx;
      s.precision = precision;
      s.in_dtype = in.dtype;
      set_layout_for(precision, in.dtype, s.tt_f64, s.tv_f64);
      s.n = pk.n;
      s.records = pk.roff.empty() ? 0 : pk.roff.back();
      s.points = pk.points;
      s.h_roff = pk.roff;
      s.h_perm = pk.perm;
      s.has_empty = std::find(pk.src.begin(), pk.src.end(), int64_t(-1)) != pk.src.end();
      use_device(ctx);
      upload(s.roff, pk.roff);
      upload(s.perm, pk.perm);
      upload(s.src, pk.src);
      upload(s.rec, interleave_records(pk, s.tt_f64, s.tv_f64));
    }

    // Device packing of points already on the device (tp / vp, strides in
    // bytes; PCF i's points offset by pbase[i] bytes when given), for a set
    // whose type fields are filled in. The lowest PCF breaking a rule is
    // re-validated on the host for the host packer's exact message.
    void pack_on_gpu(Ctx* ctx, Set& s, const std::vector<int64_t>& h_off, const unsigned char* tp,
                     const unsigned char* vp, int64_t t_stride, int64_t v_stride, int32_t dtype,
                     const std::vector<int64_t>* pbase)
    {
      const int64_t n = s.n;
      const int64_t es = dtype == SBAMD_F64 ? 8 : 4;
      const hipStream_t st = ctx->stream;
      DeviceBuffer d_off, d_pbase;
      upload(d_off, h_off);
      if (pbase)
        upload(d_pbase, *pbase);
      PackParams P;
      std::memset(&P, 0, sizeof(P));
      P.n = n;
      P.off = d_off.as<int64_t>();
      P.pbase = pbase ? d_pbase.as<int64_t>() : nullptr;
      P.t = tp;
      P.v = vp;
      P.t_stride = t_stride;
      P.v_stride = v_stride;
      P.tt_f64 = s.tt_f64;
      P.tv_f64 = s.tv_f64;
      DeviceBuffer d_m(static_cast<size_t>(n) * sizeof(int64_t));
      DeviceBuffer d_bad(sizeof(unsigned long long));
      const unsigned long long none = ~0ull;
      SBAMD_HIP(hipMemcpy(d_bad.ptr, &none, sizeof(none), hipMemcpyHostToDevice));
      P.m = d_m.as<int64_t>();
      P.bad = d_bad.as<unsigned long long>();
      launch_pack_count(P, dtype == SBAMD_F64, st);
      unsigned long long bad = none;
      std::vector<int64_t> m(static_cast<size_t>(n));
      SBAMD_HIP(hipMemcpyAsync(&bad, d_bad.ptr, sizeof(bad), hipMemcpyDeviceToHost, st));
      if (n > 0)
        SBAMD_HIP(hipMemcpyAsync(m.data(), d_m.ptr, m.size() * sizeof(int64_t), hipMemcpyDeviceToHost, st));
      SBAMD_HIP(hipStreamSynchronize(st));
      if (bad != none)
      {
        // the host rules on the first offending PCF give the message
        const int64_t i = static_cast<int64_t>(bad);
        const int64_t b = h_off[i], cnt = h_off[i + 1] - b;
        const int64_t pb = pbase ? (*pbase)[i] : 0;
        std::vector<unsigned char> raw_t(static_cast<size_t>(cnt * es)), raw_v(raw_t.size());
        for (int64_t k = 0; k < cnt; ++k)
        {
          const hipMemcpyKind kind = hipMemcpyDeviceToHost;
          SBAMD_HIP(hipMemcpy(raw_t.data() + k * es, tp + pb + (b + k) * t_stride, es, kind));
          SBAMD_HIP(hipMemcpy(raw_v.data() + k * es, vp + pb + (b + k) * v_stride, es, kind));
        }
        const int64_t one_off[2] = {0, cnt};
        sbamd_pcf_set one;
        one.n = 1;
        one.offsets = one_off;
        one.t = raw_t.data();
        one.v = raw_v.data();
        one.dtype = dtype;
        validate_set(one);
        fail(SBAMD_ERR_RUNTIME, "device validation flagged PCF " + std::to_string(i) +
                                    " but the host rules accept it");
      }

      std::vector<int64_t> perm;
      sort_by_records(m, perm);
      std::vector<int64_t> roff(static_cast<size_t>(n) + 1, 0), qinv(static_cast<size_t>(n));
      for (int64_t q = 0; q < n; ++q)
      {
        roff[q + 1] = roff[q] + m[perm[q]];
        qinv[perm[q]] = q;
      }
      s.records = roff.back();
      s.h_roff = roff;
      s.h_perm = perm;
      upload(s.roff, roff);
      upload(s.perm, perm);
      DeviceBuffer d_qinv;
      upload(d_qinv, qinv);
      s.rec = DeviceBuffer(static_cast<size_t>(s.records) * record_bytes(s.tt_f64));
      s.src = DeviceBuffer(static_cast<size_t>(s.records) * sizeof(int64_t));
      P.roff = s.roff.as<int64_t>();
      P.qinv = d_qinv.as<int64_t>();
      P.rec = s.rec.ptr;
      P.src = s.src.as<int64_t>();
      launch_pack_records(P, dtype == SBAMD_F64, st);
      SBAMD_HIP(hipStreamSynchronize(st));
    }

    // The same set as make_set, validated and packed on the device. Host
    // views are copied once as they lie (one copy when times and values
    // interleave); device views are read in place.
    void make_set_gpu(Ctx* ctx, const sbamd_pcf_view& in, int32_t precision, Set& s)
    {
      if (in.dtype != SBAMD_F32 && in.dtype != SBAMD_F64)
        fail(SBAMD_ERR_INVALID, "PCF set dtype must be SBAMD_F32 or SBAMD_F64");
      if (in.location != SBAMD_HOST && in.location != SBAMD_DEVICE)
        fail(SBAMD_ERR_INVALID, "PCF view location must be SBAMD_HOST or SBAMD_DEVICE");
      if (in.n < 0)
        fail(SBAMD_ERR_INVALID, "PCF set size must be non-negative");
      if (in.n > std::numeric_limits<int32_t>::max() / 2)
        fail(SBAMD_ERR_INVALID, "PCF set too large (more than 2^30 PCFs)");
      if (in.n > 0 && !in.offsets)
        fail(SBAMD_ERR_INVALID, "PCF set offsets are NULL");
      const int64_t es = in.dtype == SBAMD_F64 ? 8 : 4;
      use_device(ctx);

      std::vector<int64_t> h_off(static_cast<size_t>(in.n) + 1, 0);
      if (in.n > 0)
      {
        if (in.location == SBAMD_HOST)
          std::memcpy(h_off.data(), in.offsets, h_off.size() * sizeof(int64_t));
        else
          SBAMD_HIP(hipMemcpy(h_off.data(), in.offsets, h_off.size() * sizeof(int64_t),
                              hipMemcpyDeviceToHost));
      }
      // offsets: the host packer's rules
      check_set_offsets(in.n, h_off.data(), in.dtype);
      const int64_t M = in.n > 0 ? h_off[in.n] : 0;
      if (M > 0 && (!in.t || !in.v))
        fail(SBAMD_ERR_INVALID, "PCF set time/value arrays are NULL");
      if (M > 0 && (in.t_stride < es || in.v_stride < es))
        fail(SBAMD_ERR_INVALID, "PCF view strides must be at least the element size");

      s.ctx = ctx;
      s.precision = precision;
      s.in_dtype = in.dtype;
      set_layout_for(precision, in.dtype, s.tt_f64, s.tv_f64);
      s.n = in.n;
      s.points = M;
      s.has_empty = false;
      for (int64_t i = 0; i < in.n; ++i)
        s.has_empty |= h_off[i + 1] == h_off[i];

      DeviceBuffer d_pts, d_t, d_v;
      const unsigned char* tp = static_cast<const unsigned char*>(in.t);
      const unsigned char* vp = static_cast<const unsigned char*>(in.v);
      if (M > 0 && in.location == SBAMD_HOST)
      {
        // byte spans of the times and the values; one copy of their union
        // when they overlap (AoS points), else one copy each
        const unsigned char* t_end = tp + (M - 1) * in.t_stride + es;
        const unsigned char* v_end = vp + (M - 1) * in.v_stride + es;
        if (tp < v_end && vp < t_end)
        {
          const unsigned char* lo = std::min(tp, vp);
          const unsigned char* hi = std::max(t_end, v_end);
          d_pts = DeviceBuffer(static_cast<size_t>(hi - lo));
          SBAMD_HIP(hipMemcpy(d_pts.ptr, lo, d_pts.bytes, hipMemcpyHostToDevice));
          tp = d_pts.as<unsigned char>() + (tp - lo);
          vp = d_pts.as<unsigned char>() + (vp - lo);
        }
        else
        {
          d_t = DeviceBuffer(static_cast<size_t>(t_end - tp));
          d_v = DeviceBuffer(static_cast<size_t>(v_end - vp));
          SBAMD_HIP(hipMemcpy(d_t.ptr, tp, d_t.bytes, hipMemcpyHostToDevice));
          SBAMD_HIP(hipMemcpy(d_v.ptr, vp, d_v.bytes, hipMemcpyHostToDevice));
          tp = d_t.as<unsigned char>();
          vp = d_v.as<unsigned char>();
        }
      }

      pack_on_gpu(ctx, s, h_off, tp, vp, in.t_stride, in.v_stride, in.dtype, nullptr);
    }

    // A device set straight from a file image of a pcf32 / pcf64 tensor or
    // object: the payload (point counts and points as they lie) is copied to
    // the device once and PCF i's points are read at its own byte offset.
    void set_from_image(Ctx* ctx, const void* buf, uint64_t len, int32_t precision, Set& s)
    {
      sbamd_io_info info;
      io_parse(buf, len, info);
      if (info.base_format != 100)
        fail(SBAMD_ERR_INVALID, "the file does not hold PCFs with floating-point times and values "
                                "(element format (" + std::to_string(info.base_format) + ", " +
                                    std::to_string(info.sub_format) + "))");
      const uint64_t E = info.elements;
      if (E > static_cast<uint64_t>(std::numeric_limits<int32_t>::max() / 2))
        fail(SBAMD_ERR_INVALID, "PCF set too large (more than 2^30 PCFs)");
      std::vector<uint64_t> counts(E), starts(E);
      const uint64_t end = io_index(buf, len, info, counts.data(), starts.data());
      const int32_t dtype = info.sub_format == 64 ? SBAMD_F64 : SBAMD_F32;
      const int64_t es = info.sub_format / 8, pb = 2 * es;
      std::vector<int64_t> h_off(E + 1, 0), pbase(E);
      for (uint64_t e = 0; e < E; ++e)
        h_off[e + 1] = h_off[e] + static_cast<int64_t>(counts[e]);
      for (uint64_t e = 0; e < E; ++e)
        pbase[e] = static_cast<int64_t>(starts[e] - info.payload) - h_off[e] * pb;
      use_device(ctx);
      s.ctx = ctx;
      s.precision = precision;
      s.in_dtype = dtype;
      set_layout_for(precision, dtype, s.tt_f64, s.tv_f64);
      s.n = static_cast<int64_t>(E);
      s.points = h_off[E];
      s.has_empty = false;
      for (uint64_t e = 0; e < E; ++e)
        s.has_empty |= counts[e] == 0;
      DeviceBuffer d_img(static_cast<size_t>(end - info.payload));
      if (d_img.bytes)
        SBAMD_HIP(hipMemcpy(d_img.ptr, static_cast<const unsigned char*>(buf) + info.payload, d_img.bytes,
                            hipMemcpyHostToDevice));
      const unsigned char* tp = d_img.as<unsigned char>();
      pack_on_gpu(ctx, s, h_off, tp, tp + es, pb, pb, dtype, &pbase);
    }

    RowsParams rows_params(int32_t out_kind, int64_t n_rows, int64_t n_cols, const void* d, int32_t dtype)
    {
      if (out_kind != SBAMD_OUT_DISTMAT && out_kind != SBAMD_OUT_DENSE)
        fail(SBAMD_ERR_INVALID, "row reductions take SBAMD_OUT_DISTMAT or SBAMD_OUT_DENSE matrices");
      if (dtype != SBAMD_F32 && dtype != SBAMD_F64)
        fail(SBAMD_ERR_INVALID, "dtype must be SBAMD_F32 or SBAMD_F64");
      if (out_kind == SBAMD_OUT_DISTMAT)
        n_cols = n_rows;
      if (n_rows < 0 || n_cols < 0)
        fail(SBAMD_ERR_INVALID, "matrix extents must be non-negative");
      if (n_rows > std::numeric_limits<int32_t>::max() || n_cols > std::numeric_limits<int32_t>::max() - 1)
        fail(SBAMD_ERR_INVALID, "matrix extents must be below 2^31");
      const bool any = out_kind == SBAMD_OUT_DISTMAT ? n_rows > 1 : n_rows > 0 && n_cols > 0;
      if (any)
        need(d, "d_device");
      RowsParams P;
      std::memset(&P, 0, sizeof(P));
      P.d = d;
      P.n_rows = n_rows;
      P.n_cols = n_cols;
      P.compressed = out_kind == SBAMD_OUT_DISTMAT;
      P.f64 = dtype == SBAMD_F64;
      return P;
    }

    // Large host sets are packed on the GPU (the same bytes as the host
    // packer, one upload of t and v instead of a single-threaded pass).
    constexpr int64_t GPU_PACK_MIN_POINTS = int64_t(1) << 16;

    void make_set_auto(Ctx* ctx, const sbamd_pcf_set& in, int32_t precision, Set& s)
    {
      const int64_t M = in.n > 0 && in.offsets ? in.offsets[in.n] : 0;
      if (M < GPU_PACK_MIN_POINTS)
      {
        make_set(ctx, in, precision, s);
        return;
      }
      sbamd_pcf_view view;
      view.n = in.n;
      view.offsets = in.offsets;
      view.t = in.t;
      view.v = in.v;
      view.t_stride = view.v_stride = in.dtype == SBAMD_F64 ? 8 : 4;
      view.dtype = in.dtype;
      view.location = SBAMD_HOST;
      make_set_gpu(ctx, view, precision, s);
    }

    std::vector<int64_t> sorted_counts(const Set& s)
    {
      std::vector<int64_t> m(static_cast<size_t>(s.n));
      for (int64_t q = 0; q < s.n; ++q)
        m[q] = s.h_roff[q + 1] - s.h_roff[q];
      return m;
    }

    uint64_t tri(uint64_t n) { return n * (n - 1) / 2; }

    void finish_plan(Plan& pl)
    {
      use_device(pl.ctx);
      upload(pl.probs, pl.h_probs);
      upload(pl.tiles, pl.h_tiles);
      std::vector<int64_t> qtab(NCLASS * 2 * DEVICE_QUEUES);
      for (int k = 0; k < NCLASS; ++k)
        for (int x = 0; x < DEVICE_QUEUES; ++x)
        {
          qtab[k * 2 * DEVICE_QUEUES + x] = pl.xq[k][x];
          qtab[k * 2 * DEVICE_QUEUES + DEVICE_QUEUES + x] = pl.xq[k][x + 1] - pl.xq[k][x];
        }
      upload(pl.qtab, qtab);
      if (!pl.counter.ptr)
        // one tile counter per queue and class launch, 64 bytes apart
        pl.counter = DeviceBuffer(64 * NCLASS * DEVICE_QUEUES);
    }

    FwdParams fwd_params(const Plan& pl, void* out, int32_t out_dtype, double* row_sum)
    {
      FwdParams P;
      std::memset(&P, 0, sizeof(P));
      const Set* rs = pl.sets[0];
      const Set* cs = pl.cdist ? pl.sets[1] : pl.sets[0];
      if (pl.batched)
      {
        P.r_off = pl.b_roff.as<int64_t>();
        P.r_rec = pl.b_rec.ptr;
        P.r_perm = pl.b_perm.as<int64_t>();
        P.c_off = P.r_off;
        P.c_rec = P.r_rec;
        P.c_perm = P.r_perm;
      }
      else
      {
        P.r_off = rs->roff.as<int64_t>();
        P.r_rec = rs->rec.ptr;
        P.r_perm = rs->perm.as<int64_t>();
        P.c_off = cs->roff.as<int64_t>();
        P.c_rec = cs->rec.ptr;
        P.c_perm = cs->perm.as<int64_t>();
      }
      P.tiles = pl.tiles.as<Tile>();
      P.probs = pl.probs.as<Problem>();
      P.out = out;
      P.row_sum = row_sum;
      P.ntiles = static_cast<int64_t>(pl.h_tiles.size());
      P.out_cols = pl.cdist ? pl.sets[1]->n : 0;
      P.p = pl.p;
      P.out_kind = pl.out_kind;
      P.out_f64 = out_dtype == SBAMD_F64 ? 1 : 0;
      P.R = pl.R;
      P.C = pl.C;
      P.cap = pl.cap;
      P.ppl = pl.ppl;
      P.stride = pl.R + pl.C;
      P.symmetric = pl.cdist ? 0 : 1;
      P.include_diag = pl.out_kind == SBAMD_OUT_SYMMAT ? 1 : 0;
      return P;
    }

    // Work units of a plan's forward launch: pairs for the long-pair kernel,
    // tiles otherwise. The job path runs a plan in unit ranges.
    int64_t work_units(const Plan& pl)
    {
      return pl.use_long ? static_cast<int64_t>(pl.pairs) : static_cast<int64_t>(pl.h_tiles.size());
    }

    // Long-pair kernel on a plan of few, long pairs (FP32C), or when asked.
    // One lane per pair fills the device only with ~1536 pairs in flight per
    // CU; below a quarter of that, with >= 2048 segments per pair on
    // average, a wavefront per pair is the better split.
    void decide_long(Plan& pl)
    {
      if (pl.batched || pl.pairs == 0)
        pl.use_long = false;
      else if (pl.long_mode == 1)
        pl.use_long = true;
      else if (pl.long_mode == 0)
        pl.use_long = false;
      else
        pl.use_long = pl.precision == SBAMD_PREC_FP32C &&
                      pl.pairs * 4 <= static_cast<uint64_t>(pl.ctx->cu_count) * 1536 &&
                      pl.segments >= pl.pairs * 2048;
      pl.bump();
    }

    // A batched plan reads its own copy of the member sets' records: copy
    // again (on `stream`, into the same buffer) the members whose values were
    // updated since.
    void refresh_batch(Plan& pl, hipStream_t stream)
    {
      const size_t rb = static_cast<size_t>(record_bytes(pl.tt_f64));
      for (size_t b = 0; b < pl.sets.size(); ++b)
      {
        const Set* s = pl.sets[b];
        if (s->vgen == pl.b_vgen[b])
          continue;
        if (s->records > 0)
          SBAMD_HIP(hipMemcpyAsync(static_cast<char*>(pl.b_rec.ptr) + static_cast<size_t>(pl.b_at[b]) * rb,
                                   s->rec.ptr, static_cast<size_t>(s->records) * rb,
                                   hipMemcpyDeviceToDevice, stream));
        pl.b_vgen[b] = s->vgen;
      }
    }

    void execute_range(const Plan& pl, void* out, int32_t out_dtype, double* row_sum, int64_t t0,
                       int64_t t1, hipStream_t stream)
    {
      if (pl.use_long)
      {
        const FwdParams F = fwd_params(pl, out, out_dtype, row_sum);
        LongParams L;
        std::memset(&L, 0, sizeof(L));
        L.r_off = F.r_off;
        L.r_rec = F.r_rec;
        L.r_perm = F.r_perm;
        L.c_off = F.c_off;
        L.c_rec = F.c_rec;
        L.c_perm = F.c_perm;
        L.out = out;
        L.row_sum = row_sum;
        L.q0 = t0;
        L.q1 = t1;
        L.n_cols = pl.cdist ? pl.sets[1]->n : 0;
        L.p = pl.p;
        L.out_kind = pl.out_kind;
        L.out_f64 = out_dtype == SBAMD_F64 ? 1 : 0;
        launch_long(L, pl.tt_f64, pl.tv_f64, kernel_op(pl.op, pl.p),
                    pl.precision == SBAMD_PREC_FP32C ? 1 : 0, pl.ctx->cu_count, stream);
        return;
      }
      if (pl.persistent && t1 > t0)
        // reset by a kernel, not hipMemsetAsync, so graph replays stay correct
        launch_fill_zero(pl.counter.ptr, pl.counter.bytes, stream);
      // one launch per length class (a single one without per-class
      // shapes), the classes of longer PCFs first
      for (int k = NCLASS - 1; k >= 0; --k)
      {
        const int64_t a = std::max(t0, pl.class_begin[k]), b = std::min(t1, pl.class_begin[k + 1]);
        if (a >= b)
          continue;
        const Shape sh = pl.shape_of(k);
        if (fwd_lds_bytes(pl.tt_f64, pl.tv_f64, sh.R, sh.C, sh.cap) > pl.ctx->lds_per_block)
          fail(SBAMD_ERR_INVALID, "tile shape needs more LDS than a workgroup may use");
        FwdParams P = fwd_params(pl, out, out_dtype, row_sum);
        P.R = sh.R;
        P.C = sh.C;
        P.cap = sh.cap;
        P.ppl = sh.ppl;
        P.stride = sh.R + sh.C;
        P.tiles = pl.tiles.as<Tile>() + a;
        P.ntiles = b - a;
        if (pl.persistent)
        {
          P.work_counter = reinterpret_cast<uint32_t*>(pl.counter.as<unsigned char>() + 64 * DEVICE_QUEUES * k);
          // the class's queues when the whole class runs, else one queue
          const bool whole = a == pl.class_begin[k] && b == pl.class_begin[k + 1];
          P.q_tab = whole ? pl.qtab.as<int64_t>() + 2 * DEVICE_QUEUES * k : nullptr;
        }
        launch_forward(P, pl.tt_f64, pl.tv_f64, kernel_op(pl.op, pl.p),
                       pl.precision == SBAMD_PREC_FP32C ? 1 : 0, stream);
      }
    }

    void check_out_dtype(int32_t d)
    {
      if (d != SBAMD_F32 && d != SBAMD_F64)
        fail(SBAMD_ERR_INVALID, "output dtype must be SBAMD_F32 or SBAMD_F64");
    }

    uint64_t tile_pairs(const Tile& t, bool include_diag)
    {
      if (t.flags & TILE_DIAG)
      {
        // pairs (r0 + a, c0 + b) with c0 + b <= r0 + a (- 1 without the diagonal)
        uint64_t n = 0;
        for (int a = 0; a < t.nr; ++a)
        {
          const int64_t upto = static_cast<int64_t>(t.r0) + a - (include_diag ? 0 : 1) - t.c0 + 1;
          n += static_cast<uint64_t>(std::min<int64_t>(std::max<int64_t>(upto, 0), t.nc));
        }
        return n;
      }
      return static_cast<uint64_t>(t.nr) * static_cast<uint64_t>(t.nc);
    }

    // Working arrays of one backward run (kernels_bwd.hip).
    struct BwdWork
    {
      double* key0;
      double* val0;
      int32_t* col0;
      int64_t* src0;
      double* keys[2];
      int32_t* ids[2];
      BwdEvent* events;
      BwdEvent* events2;
      double* q;
      double* a;
      double* w;
      int32_t* ties; // cross-tie flag of the lists (NULL: not kept, ties assumed)
    };

    int backward_checks(const Plan* plan, const void* grad_out_device, const void* fwd_out_device,
                        int32_t out_dtype, void* grad_v_device, void* grad_t_device, int32_t grad_dtype)
    {
      need(plan, "plan");
      need(grad_v_device, "grad_v_device");
      need(grad_t_device, "grad_t_device");
      check_out_dtype(out_dtype);
      check_out_dtype(grad_dtype);
      if (plan->cdist || plan->batched)
        fail(SBAMD_ERR_INVALID, "backward supports single-set pdist / l2_kernel plans");
      const int kop = kernel_op(plan->op, plan->p);
      if (!((kop == OPK_L2 && plan->out_kind == SBAMD_OUT_DISTMAT) ||
            (kop == OPK_DOT && plan->out_kind == SBAMD_OUT_SYMMAT)))
        fail(SBAMD_ERR_INVALID, "backward supports L2 pdist (p = 2) and the L2 kernel");
      if (!plan->tt_f64)
        fail(SBAMD_ERR_INVALID, "backward needs a float64 PCF set");
      if (plan->out_count > 0)
      {
        need(grad_out_device, "grad_out_device");
        if (kop == OPK_L2)
          need(fwd_out_device, "fwd_out_device");
      }
      if (plan->sets[0]->records > INT32_MAX)
        fail(SBAMD_ERR_INVALID, "backward supports sets of up to 2^31 - 1 records");
      return kop;
    }

    // The backward launches on working arrays `wk` (no allocation here).
    void run_backward(const Plan& pl, int kop, const BwdWork& wk, bool lists_ready, const void* grad_out_device,
                      const void* fwd_out_device, int32_t out_dtype, void* grad_v_device,
                      void* grad_t_device, int32_t grad_dtype, hipStream_t st)
    {
      const Set* s = pl.sets[0];
      const size_t gb = grad_dtype == SBAMD_F64 ? 8 : 4;
      const size_t M = static_cast<size_t>(s->records);
      const size_t N = static_cast<size_t>(s->n);
      {
        // the gradients and the backward's accumulators, cleared by one launch
        void* const ptrs[5] = {grad_v_device, grad_t_device, wk.q, wk.a, wk.w};
        const size_t bytes[5] = {static_cast<size_t>(s->points) * gb, static_cast<size_t>(s->points) * gb,
                                 M * 8, M * 8, N * 8};
        launch_fill_zero_multi(ptrs, bytes, 5, st);
      }
      BwdParams B;
      std::memset(&B, 0, sizeof(B));
      B.n = s->n;
      B.records = s->records;
      B.op = kop;
      B.in_f64 = out_dtype == SBAMD_F64 ? 1 : 0;
      B.w_f64 = pl.tv_f64; // FP64 sets keep FP64 pair weights
      B.val_f64 = pl.tv_f64;
      B.grad_f64 = grad_dtype == SBAMD_F64 ? 1 : 0;
      // chunks of 64 partners with FP32 pair weights, 32 with FP64: a lane
      // keeps its row's weights of the chunk in 64 vector registers
      B.ch = pl.tv_f64 ? BWD_GROUP / 2 : BWD_GROUP;
      B.roff = s->roff.as<int64_t>();
      B.rec = s->rec.ptr;
      B.perm = s->perm.as<int64_t>();
      B.src = s->src.as<int64_t>();
      B.ooff = pl.bwd_ooff.as<int64_t>();
      B.key0 = wk.key0;
      B.val0 = wk.val0;
      B.col0 = wk.col0;
      B.src0 = wk.src0;
      for (int b = 0; b < 2; ++b)
      {
        B.keys[b] = wk.keys[b];
        B.ids[b] = wk.ids[b];
      }
      B.events = wk.events;
      B.events2 = wk.events2;
      B.grad_out = grad_out_device;
      B.fwd_out = fwd_out_device;
      B.q = wk.q;
      B.a = wk.a;
      B.wsum = wk.w;
      B.grad_v = grad_v_device;
      B.grad_t = grad_t_device;
      B.cross_ties = wk.ties;
      launch_backward(B, lists_ready, st);
    }

    // Record offsets of a single-set plan's PCFs in the original order, for
    // the backward's position layout.
    void upload_ooff(Plan& pl)
    {
      const Set* s = pl.sets[0];
      const size_t N = static_cast<size_t>(s->n);
      std::vector<int64_t> ooff(N + 1, 0);
      for (size_t q = 0; q < N; ++q)
        ooff[static_cast<size_t>(s->h_perm[q]) + 1] = s->h_roff[q + 1] - s->h_roff[q];
      for (size_t i = 0; i < N; ++i)
        ooff[i + 1] += ooff[i];
      upload(pl.bwd_ooff, ooff);
    }

    // A plan over rows (and cols: a cross plan). Cross plans normally write a
    // dense matrix; the host-to-host jobs also make cross plans that write
    // compressed outputs (pairs whose row index exceeds every column index).
    // out_base is added to every output index.
    // Default tile: 16 rows x 32 columns, 64 staged records per slot (48
    // slots: three workgroups of 8 wavefronts per CU, set by LDS). Measured
    // on MI355X: fewer phases per tile than 16 x 16 x 48 at the same
    // occupancy, for pdists (m ~ 50..500) and long cdists (m ~ 500..5000)
    // alike; 16 x 16 x 48 where a workgroup cannot hold that much LDS.
    void set_default_shape(Plan& pl)
    {
      pl.R = 16;
      pl.C = 32;
      pl.cap = 64;
      pl.ppl = 1;
      if (fwd_lds_bytes(pl.tt_f64, pl.tv_f64, pl.R, pl.C, pl.cap) > pl.ctx->lds_per_block)
      {
        pl.R = pl.C = 16;
        pl.cap = 48;
      }
    }

    std::unique_ptr<sbamd_plan> new_plan(Ctx* ctx, const Set* rows, const Set* cols, int32_t op, double p,
                                         int32_t out_kind, const sbamd_tile_shape* shape, int64_t out_base)
    {
      if (op == SBAMD_OP_LP && !(p >= 1.0))
        fail(SBAMD_ERR_INVALID, "p must be >= 1.");
      (void)kernel_op(op, p);
      auto pl = std::make_unique<sbamd_plan>();
      pl->ctx = ctx;
      pl->op = op;
      pl->p = op == SBAMD_OP_LP ? p : 1.0;
      pl->out_kind = out_kind;
      pl->precision = rows->precision;
      pl->tt_f64 = rows->tt_f64;
      pl->tv_f64 = rows->tv_f64;
      pl->sets.push_back(rows);
      if (cols)
      {
        if (cols->precision != rows->precision || cols->tt_f64 != rows->tt_f64 ||
            cols->tv_f64 != rows->tv_f64)
          fail(SBAMD_ERR_INVALID, "row and column sets must share dtype and precision");
        pl->cdist = true;
        pl->sets.push_back(cols);
        pl->pairs = static_cast<uint64_t>(rows->n) * static_cast<uint64_t>(cols->n);
        pl->out_count = pl->pairs;
        pl->segments = static_cast<uint64_t>(cols->n) * rows->records +
                       static_cast<uint64_t>(rows->n) * cols->records - pl->pairs;
      }
      else
      {
        const uint64_t n = static_cast<uint64_t>(rows->n);
        const uint64_t M = static_cast<uint64_t>(rows->records);
        if (out_kind == SBAMD_OUT_DISTMAT)
        {
          pl->pairs = n > 1 ? n * (n - 1) / 2 : 0;
          pl->segments = n > 1 ? (n - 1) * M - pl->pairs : 0;
        }
        else if (out_kind == SBAMD_OUT_SYMMAT)
        {
          pl->pairs = n * (n + 1) / 2;
          pl->segments = n > 0 ? n * M - (n > 1 ? n * (n - 1) / 2 : 0) : 0;
        }
        else
          fail(SBAMD_ERR_INVALID, "pdist / kernel plans need SBAMD_OUT_DISTMAT or SBAMD_OUT_SYMMAT");
        pl->out_count = pl->pairs;
      }
      Problem pr{0, 0, out_base, rows->n};
      pl->h_probs.push_back(pr);
      pl->input_bytes = 0;
      for (const Set* st : pl->sets)
        pl->input_bytes += st->rec.bytes + st->roff.bytes;
      set_default_shape(*pl);
      if (shape)
      {
        if (shape->rows)
          pl->R = shape->rows;
        if (shape->cols)
          pl->C = shape->cols;
        else if (shape->rows && !pl->cdist)
          pl->C = pl->R; // only a side given for a symmetric plan: a square tile
        if (shape->cap)
          pl->cap = shape->cap;
        if (shape->lane_pairs)
          pl->ppl = shape->lane_pairs;
      }
      check_shape(pl->R, pl->C, pl->cap, pl->ppl);
      pl->rebuild_tiles();
      if (out_base != 0)
        pl->long_mode = 0; // the long-pair kernel writes whole outputs only
      decide_long(*pl);
      if (!pl->cdist && out_base == 0)
        upload_ooff(*pl);
      return pl;
    }

    // Autotuner candidates: 32 x 32, 16 x 16 and (plans too small to fill
    // the device with 16 x 16 tiles) 8 x 8 tiles, one or two pairs per lane,
    // LDS windows that leave room for two or more workgroups per CU.
    std::vector<Shape> tune_candidates(const Plan& plan)
    {
      std::vector<Shape> cands;
      const int sides[] = {32, 16, 8};
      const int caps[] = {32, 48, 64, 96};
      const size_t budget = std::min<size_t>(plan.ctx->lds_per_block, 80 * 1024);
      uint64_t mid_tiles = 0; // 16 x 16 tiles of the plan
      for (size_t b = 0; b < plan.sets.size(); ++b)
      {
        const uint64_t nr = (static_cast<uint64_t>(plan.sets[b]->n) + 15) / 16;
        if (plan.cdist)
        {
          mid_tiles = nr * ((static_cast<uint64_t>(plan.sets[1]->n) + 15) / 16);
          break;
        }
        mid_tiles += nr * (nr + 1) / 2;
      }
      const bool small = mid_tiles < 4ull * static_cast<uint64_t>(plan.ctx->cu_count);
      for (int R : sides)
        for (int C : sides)
        {
          if ((R < 16 || C < 16) && !small)
            continue;
          for (int ppl = 1; ppl <= 2; ++ppl)
          {
            if (!tile_shape_ok(R, C, 2, ppl))
              continue;
            for (int cap : caps)
              if (fwd_lds_bytes(plan.tt_f64, plan.tv_f64, R, C, cap) <= budget)
                cands.push_back(Shape{R, C, cap, ppl});
          }
        }
      return cands;
    }

    // Cache of tuned shapes, keyed by device architecture, operation,
    // record layout, plan kind, length class and a coarse histogram of that
    // class (log2 buckets of its PCF count, mean record count and pair
    // count). Kept for the process; with SBAMD_TUNE_CACHE naming a file it
    // is also read from and appended to that file (one "key R C cap ppl"
    // line per entry).
    std::mutex g_tune_mu;
    std::unordered_map<std::string, Shape>& tune_cache()
    {
      static std::unordered_map<std::string, Shape> m;
      static bool loaded = false;
      if (!loaded)
      {
        loaded = true;
        if (const char* path = std::getenv("SBAMD_TUNE_CACHE"))
          if (FILE* f = std::fopen(path, "r"))
          {
            char key[512];
            Shape sh;
            while (std::fscanf(f, "%511s %d %d %d %d", key, &sh.R, &sh.C, &sh.cap, &sh.ppl) == 5)
              if (tile_shape_ok(sh.R, sh.C, sh.cap, sh.ppl))
                m[key] = sh;
            std::fclose(f);
          }
      }
      return m;
    }

    std::atomic<uint64_t> g_tune_hits{0}, g_tune_misses{0};

    bool tune_cache_get(const std::string& key, Shape& sh)
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto& m = tune_cache();
      auto it = m.find(key);
      if (it == m.end())
      {
        ++g_tune_misses;
        return false;
      }
      ++g_tune_hits;
      sh = it->second;
      return true;
    }

    void tune_cache_put(const std::string& key, const Shape& sh)
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      tune_cache()[key] = sh;
      if (const char* path = std::getenv("SBAMD_TUNE_CACHE"))
        if (FILE* f = std::fopen(path, "a"))
        {
          std::fprintf(f, "%s %d %d %d %d\n", key.c_str(), sh.R, sh.C, sh.cap, sh.ppl);
          std::fclose(f);
        }
    }

    int log2_bucket(uint64_t x)
    {
      int b = 0;
      while (x > 1)
      {
        x >>= 1;
        ++b;
      }
      return b;
    }

    std::string tune_key(const Plan& pl, int k)
    {
      uint64_t pcfs = 0, recs = 0;
      for (size_t b = 0; b < (pl.cdist ? 1 : pl.sets.size()); ++b)
      {
        const Set& s = *pl.sets[b];
        for (int64_t q = 0; q < s.n; ++q)
        {
          const int64_t m = s.h_roff[q + 1] - s.h_roff[q];
          if (length_class(m) == k)
          {
            ++pcfs;
            recs += static_cast<uint64_t>(m);
          }
        }
      }
      // pairs whose longer PCF is of class k (independent of the tile shapes)
      uint64_t pairs = 0;
      if (pl.cdist)
      {
        uint64_t r_k = 0, r_lt = 0, c_k = 0, c_le = 0;
        for (int64_t q = 0; q < pl.sets[0]->n; ++q)
        {
          const int c = length_class(pl.sets[0]->h_roff[q + 1] - pl.sets[0]->h_roff[q]);
          r_k += c == k;
          r_lt += c < k;
        }
        for (int64_t q = 0; q < pl.sets[1]->n; ++q)
        {
          const int c = length_class(pl.sets[1]->h_roff[q + 1] - pl.sets[1]->h_roff[q]);
          c_k += c == k;
          c_le += c <= k;
        }
        pairs = r_k * c_le + r_lt * c_k;
      }
      else
        for (const Set* st : pl.sets)
        {
          uint64_t below = 0, in = 0;
          for (int64_t q = 0; q < st->n; ++q)
          {
            const int c = length_class(st->h_roff[q + 1] - st->h_roff[q]);
            below += c < k;
            in += c == k;
          }
          pairs += in * below + in * (in + 1) / 2;
        }
      std::string key = pl.ctx->arch;
      for (char& c : key)
        if (c == ' ')
          c = '_';
      key += "|op" + std::to_string(kernel_op(pl.op, pl.p)) + "|t" + std::to_string(pl.tt_f64) + "v" +
             std::to_string(pl.tv_f64) + "|" + (pl.cdist ? "cdist" : (pl.batched ? "batched" : "pdist")) +
             "|k" + std::to_string(k) + "|n" + std::to_string(log2_bucket(pcfs)) + "|m" +
             std::to_string(log2_bucket(pcfs ? recs / pcfs : 0)) + "|P" + std::to_string(log2_bucket(pairs));
      return key;
    }

    // Subset of a device set: the PCFs at sorted positions `sel` (ascending,
    // so still sorted by record count), records gathered on the device. For
    // the forward only (no source points).
    void make_subset(const Set& full, const std::vector<int64_t>& sel, Set& s, DeviceBuffer& d_sel,
                     hipStream_t st)
    {
      s.ctx = full.ctx;
      s.precision = full.precision;
      s.in_dtype = full.in_dtype;
      s.tt_f64 = full.tt_f64;
      s.tv_f64 = full.tv_f64;
      s.n = static_cast<int64_t>(sel.size());
      s.h_roff.assign(sel.size() + 1, 0);
      s.h_perm.resize(sel.size());
      for (size_t k = 0; k < sel.size(); ++k)
      {
        const int64_t q = sel[k];
        s.h_roff[k + 1] = s.h_roff[k] + (full.h_roff[q + 1] - full.h_roff[q]);
        s.h_perm[k] = full.h_perm[q];
      }
      s.records = s.h_roff.back();
      s.points = 0;
      upload(s.roff, s.h_roff);
      upload(s.perm, s.h_perm);
      upload(d_sel, sel); // kept by the caller until the gather has run
      const int rb = record_bytes(s.tt_f64);
      s.rec = DeviceBuffer(static_cast<size_t>(s.records) * rb);
      launch_gather_subset(full.roff.as<int64_t>(), full.rec.ptr, rb, d_sel.as<int64_t>(), s.roff.as<int64_t>(),
                           s.n, s.records, s.rec.ptr, st);
    }
  }

  void Plan::bump()
  {
    gen = ++g_plan_gen;
    std::lock_guard<std::mutex> lk(g_plans_mu);
    live_plans()[this] = gen;
  }

  void Plan::rebuild_tiles()
  {
    h_tiles.clear();
    const bool diag = out_kind == SBAMD_OUT_SYMMAT;
    // Longest tiles first over [b0, b1) of h_tiles (LPT order for a launch
    // whose tiles come from several problems or regions); returns each
    // tile's work (pairs x longest pair) in the new order.
    auto lpt = [&](size_t b0, size_t b1) {
      std::vector<std::pair<int64_t, size_t>> cost(b1 - b0);
      std::vector<int64_t> work_of(b1 - b0);
      for (size_t k = b0; k < b1; ++k)
      {
        const Tile& t = h_tiles[k];
        const Set& rs = *sets[cdist ? 0 : t.prob];
        const Set& cs = *sets[cdist ? 1 : t.prob];
        int64_t mr = 0, mc = 0;
        for (int q = 0; q < t.nr; ++q)
          mr = std::max(mr, rs.h_roff[t.r0 + q + 1] - rs.h_roff[t.r0 + q]);
        for (int q = 0; q < t.nc; ++q)
          mc = std::max(mc, cs.h_roff[t.c0 + q + 1] - cs.h_roff[t.c0 + q]);
        cost[k - b0] = {-(mr + mc), k};
        work_of[k - b0] = static_cast<int64_t>(t.nr) * t.nc * (mr + mc);
      }
      std::stable_sort(cost.begin(), cost.end());
      std::vector<Tile> sorted(b1 - b0);
      std::vector<int64_t> work(b1 - b0);
      for (size_t k = 0; k < cost.size(); ++k)
      {
        sorted[k] = h_tiles[cost[k].second];
        work[k] = work_of[cost[k].second - b0];
      }
      std::copy(sorted.begin(), sorted.end(), h_tiles.begin() + static_cast<int64_t>(b0));
      return work;
    };
    // class k over [b0, b1): per-XCD queues for a batched (persistent) plan,
    // else every tile in queue 0
    auto queues = [&](int k, size_t b0, size_t b1, int qr, int qc, const std::vector<int64_t>* work) {
      if (work)
        xcd_queue_order(h_tiles, b0, b1, qr, qc, DEVICE_QUEUES, *work, xq[k]);
      else
        for (int x = 0; x <= DEVICE_QUEUES; ++x)
          xq[k][x] = x == 0 ? 0 : static_cast<int64_t>(b1 - b0);
    };
    // sorted positions where each length class starts
    auto class_bounds = [](const std::vector<int64_t>& m, int64_t* at) {
      for (int k = 0; k <= NCLASS; ++k)
        at[k] = static_cast<int64_t>(std::lower_bound(m.begin(), m.end(), CLASS_MIN[k]) - m.begin());
      at[NCLASS] = static_cast<int64_t>(m.size());
    };
    if (!per_class)
    {
      if (cdist)
      {
        auto mr = sorted_counts(*sets[0]);
        auto mc = sorted_counts(*sets[1]);
        build_tiles(mr, &mc, R, C, false, 0, fold_threshold, h_tiles);
      }
      else
        for (size_t b = 0; b < sets.size(); ++b)
        {
          auto m = sorted_counts(*sets[b]);
          build_tiles(m, nullptr, R, C, diag, static_cast<uint16_t>(b), fold_threshold, h_tiles);
        }
      if (batched)
      {
        const std::vector<int64_t> work = lpt(0, h_tiles.size());
        queues(0, 0, h_tiles.size(), R, C, &work);
      }
      else
      {
        xcd_supertile_order(h_tiles, 0, h_tiles.size(), R, C, DEVICE_XCDS);
        queues(0, 0, h_tiles.size(), R, C, nullptr);
      }
      class_begin[0] = 0;
      for (int k = 1; k <= NCLASS; ++k)
        class_begin[k] = static_cast<int64_t>(h_tiles.size());
      for (int k = 1; k < NCLASS; ++k)
        queues(k, 0, 0, R, C, nullptr);
    }
    else
    {
      // Class k's pairs are those whose longer PCF is of class k: rows of
      // class k against every shorter-or-equal column, and (cdist) rows of
      // shorter classes against columns of class k.
      std::vector<std::vector<int64_t>> ms;
      std::vector<std::array<int64_t, NCLASS + 1>> at;
      for (const Set* st : sets)
      {
        ms.push_back(sorted_counts(*st));
        std::array<int64_t, NCLASS + 1> a{};
        class_bounds(ms.back(), a.data());
        at.push_back(a);
      }
      class_begin[0] = 0;
      for (int k = 0; k < NCLASS; ++k)
      {
        const Shape sh = cls[k];
        const size_t b0 = h_tiles.size();
        if (cdist)
        {
          const auto &ar = at[0], &ac = at[1];
          build_tiles_region(ms[0], ms[1], ar[k], ar[k + 1], 0, ac[k + 1], sh.R, sh.C, false, false, 0,
                             fold_threshold, h_tiles);
          build_tiles_region(ms[0], ms[1], 0, ar[k], ac[k], ac[k + 1], sh.R, sh.C, false, false, 0,
                             fold_threshold, h_tiles);
        }
        else
          for (size_t b = 0; b < sets.size(); ++b)
          {
            const auto& a = at[b];
            const uint16_t prob = static_cast<uint16_t>(b);
            build_tiles_region(ms[b], ms[b], a[k], a[k + 1], 0, a[k], sh.R, sh.C, false, diag, prob,
                               fold_threshold, h_tiles);
            build_tiles_region(ms[b], ms[b], a[k], a[k + 1], a[k], a[k + 1], sh.R, sh.C, true, diag, prob,
                               fold_threshold, h_tiles);
          }
        const std::vector<int64_t> work = lpt(b0, h_tiles.size());
        if (!batched)
          xcd_supertile_order(h_tiles, b0, h_tiles.size(), sh.R, sh.C, DEVICE_XCDS);
        queues(k, b0, h_tiles.size(), sh.R, sh.C, batched ? &work : nullptr);
        class_begin[k + 1] = static_cast<int64_t>(h_tiles.size());
      }
    }
    finish_plan(*this);
    bump();
  }
}

// ---------------------------------------------------------------------------
// Jobs
// ---------------------------------------------------------------------------
struct sbamd_job
{
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv;
  bool done = false;
  int status = SBAMD_OK;
  std::string error;
  std::atomic<uint64_t> progress{0};
  std::atomic<uint64_t> total{0};
  std::atomic<bool> cancel{false};
};

struct sbamd_graph
{
  Ctx* ctx = nullptr;
  sbamd_plan* plan = nullptr; // the plan captured, valid while its generation is `gen`
  uint64_t gen = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

struct sbamd_plain
{
  Ctx* ctx = nullptr;
  int64_t n = 0;
  DeviceBuffer off, t, v;
};

namespace
{
  // Host-to-host job bodies (the reference's CudaPairwiseIntegrationTask /
  // CudaCrossIntegrationTask, pcf_block_op.cuh:127-286, with the block
  // pipeline of cuda_block_executor.cuh:224-297).
  constexpr uint64_t JOB_SINGLE_MAX_BYTES = uint64_t(32) << 20;
  // Banded jobs: device / pinned buffers in rotation, and the fewest bands
  constexpr int JOB_BUFFERS = 3;
  constexpr uint64_t JOB_BANDS = 8;

  template <typename T, typename D>
  std::unique_ptr<typename std::remove_pointer<T>::type, D> guard(T p, D d)
  {
    return std::unique_ptr<typename std::remove_pointer<T>::type, D>(p, d);
  }

  // memcpy over several host threads: a band's copy out of pinned memory
  // into the caller's (often freshly allocated, not yet touched) array is
  // bound by page faults and one core's bandwidth otherwise
  void parallel_copy(void* dst, const void* src, size_t bytes)
  {
    constexpr size_t MIN_PART = size_t(16) << 20;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t parts = std::min<size_t>(std::min<size_t>(hw, 16), std::max<size_t>(1, bytes / MIN_PART));
    if (parts <= 1)
    {
      std::memcpy(dst, src, bytes);
      return;
    }
    std::vector<std::thread> th;
    const size_t step = (bytes + parts - 1) / parts;
    for (size_t k = 1; k < parts; ++k)
    {
      const size_t a = k * step, b = std::min(bytes, a + step);
      if (a < b)
        th.emplace_back([=]() {
          std::memcpy(static_cast<unsigned char*>(dst) + a, static_cast<const unsigned char*>(src) + a, b - a);
        });
    }
    std::memcpy(dst, src, std::min(bytes, step));
    for (auto& t : th)
      t.join();
  }

  sbamd_tile_shape job_shape(const Ctx* ctx)
  {
    sbamd_tile_shape sh;
    sh.rows = ctx->job_R;
    sh.cols = ctx->job_C;
    sh.cap = ctx->job_cap;
    sh.lane_pairs = ctx->job_ppl;
    return sh;
  }

  // Whole output on the device (outputs that fit easily), one download at
  // the end. The tiles run in 64 ranges with an event after each; up to
  // JOB_AHEAD ranges are queued beyond the one the host waits for, so the
  // device does not idle while the host counts progress and checks for
  // cancellation between ranges.
  void job_single(Ctx* ctx, sbamd_job* jp, Set& rs, Set* cs, int32_t op, double p, int32_t out_kind,
                  void* out_host, int32_t out_dtype)
  {
    constexpr int64_t JOB_AHEAD = 3;
    const sbamd_tile_shape sh = job_shape(ctx);
    auto plan = new_plan(ctx, &rs, cs, op, p, out_kind, &sh, 0);
    const size_t eb = out_dtype == SBAMD_F64 ? 8 : 4;
    DeviceBuffer dout(static_cast<size_t>(plan->out_count) * eb);
    hipStream_t st;
    SBAMD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    auto sguard = guard(st, &hipStreamDestroy);
    const int64_t nt = work_units(*plan);
    const int64_t chunks = std::min<int64_t>(nt, 64);
    const bool diag = out_kind == SBAMD_OUT_SYMMAT;
    std::array<hipEvent_t, JOB_AHEAD + 1> ev{};
    auto ev_free = [&]() {
      for (auto& e : ev)
        if (e)
          (void)hipEventDestroy(e);
    };
    auto range = [&](int64_t c) { return std::make_pair(nt * c / chunks, nt * (c + 1) / chunks); };
    try
    {
      for (auto& e : ev)
        SBAMD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      int64_t queued = 0;
      for (int64_t c = 0; c < chunks; ++c)
      {
        for (; queued < chunks && queued <= c + JOB_AHEAD; ++queued)
        {
          const auto r = range(queued);
          execute_range(*plan, dout.ptr, out_dtype, nullptr, r.first, r.second, st);
          SBAMD_HIP(hipEventRecord(ev[static_cast<size_t>(queued % (JOB_AHEAD + 1))], st));
        }
        hipEvent_t e = ev[static_cast<size_t>(c % (JOB_AHEAD + 1))];
        while (hipEventQuery(e) == hipErrorNotReady)
        {
          if (jp->cancel.load())
            break;
          std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
        if (jp->cancel.load())
          fail(SBAMD_ERR_CANCELLED, "cancelled");
        SBAMD_HIP(hipEventSynchronize(e));
        const auto r = range(c);
        uint64_t done = 0;
        if (plan->use_long)
          done = static_cast<uint64_t>(r.second - r.first);
        else
          for (int64_t k = r.first; k < r.second; ++k)
            done += tile_pairs(plan->h_tiles[k], diag);
        jp->progress += done;
      }
      SBAMD_HIP(hipStreamSynchronize(st));
      SBAMD_HIP(hipMemcpy(out_host, dout.ptr, static_cast<size_t>(plan->out_count) * eb, hipMemcpyDeviceToHost));
    }
    catch (...)
    {
      (void)hipStreamSynchronize(st);
      ev_free();
      throw;
    }
    ev_free();
  }

  // Row bands in the original order, each writing one contiguous range of
  // the output: rows [r0, r1) against every earlier PCF (a cross plan) and
  // among themselves (a pdist / kernel plan) for the compressed layouts, or
  // against every column PCF for a dense cdist. Bands rotate over
  // JOB_BUFFERS device buffers and as many pinned host buffers (kept by the
  // context between jobs): while the device computes the next bands, a
  // band's download runs on its own stream and the host copies the band
  // before it into out_host. Each band's sets and plans are made on the host
  // while the device works on the bands before it.
  void job_bands(Ctx* ctx, sbamd_job* jp, Set& rs, Set* cs, int32_t op, double p, int32_t out_kind,
                 void* out_host, int32_t out_dtype, uint64_t chunk_elems)
  {
    const size_t eb = out_dtype == SBAMD_F64 ? 8 : 4;
    const int64_t n = rs.n;
    const int64_t ncols = cs ? cs->n : 0;
    // output elements of rows [0, r) in the original order
    auto cum = [&](int64_t r) -> uint64_t {
      const uint64_t u = static_cast<uint64_t>(r);
      if (cs)
        return u * static_cast<uint64_t>(ncols);
      return out_kind == SBAMD_OUT_SYMMAT ? u * (u + 1) / 2 : (u > 0 ? u * (u - 1) / 2 : 0);
    };
    // Bands of near-equal output, at least JOB_BANDS of them and enough that
    // none exceeds chunk_elems: band k ends at the first row whose prefix
    // reaches k / nb of the total, so a band holds at most total / nb plus
    // one row.
    const uint64_t total = cum(n);
    const uint64_t widest = cum(n) - cum(n - 1);
    if (widest > chunk_elems)
      fail(SBAMD_ERR_NOMEM, "one output row needs more device memory than the output budget allows");
    const uint64_t per = chunk_elems - widest;
    const int64_t nb = per == 0 ? n
                                : static_cast<int64_t>(std::min<uint64_t>(
                                      std::max<uint64_t>(JOB_BANDS, (total + per - 1) / per), static_cast<uint64_t>(n)));
    std::vector<int64_t> bounds(static_cast<size_t>(nb) + 1, 0);
    for (int64_t k = 1; k <= nb; ++k)
    {
      const uint64_t target = k == nb ? total : static_cast<uint64_t>(
                                                    static_cast<unsigned __int128>(total) * k / nb);
      int64_t lo = bounds[k - 1], hi = n; // first r with cum(r) >= target
      while (lo < hi)
      {
        const int64_t mid = lo + (hi - lo) / 2;
        if (cum(mid) >= target)
          hi = mid;
        else
          lo = mid + 1;
      }
      bounds[k] = lo;
    }
    bounds.erase(std::unique(bounds.begin(), bounds.end()), bounds.end());
    const int64_t nbands = static_cast<int64_t>(bounds.size()) - 1;
    for (int64_t k = 0; k < nbands; ++k)
      if (cum(bounds[k + 1]) - cum(bounds[k]) > chunk_elems)
        fail(SBAMD_ERR_RUNTIME, "band split exceeds the output budget");
    const sbamd_tile_shape sh = job_shape(ctx);

    hipStream_t st, cst;
    SBAMD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    auto g1 = guard(st, &hipStreamDestroy);
    SBAMD_HIP(hipStreamCreateWithFlags(&cst, hipStreamNonBlocking));
    auto g2 = guard(cst, &hipStreamDestroy);
    uint64_t widest_band = 0;
    for (int64_t k = 0; k < nbands; ++k)
      widest_band = std::max<uint64_t>(widest_band, cum(bounds[k + 1]) - cum(bounds[k]));
    const size_t buf_bytes = static_cast<size_t>(widest_band) * eb;
    const int nbuf = static_cast<int>(std::min<int64_t>(JOB_BUFFERS, nbands));
    std::vector<DeviceBuffer> dbuf;
    std::vector<std::unique_ptr<PinnedLease>> hbuf;
    for (int b = 0; b < nbuf; ++b)
    {
      dbuf.emplace_back(buf_bytes);
      hbuf.push_back(std::make_unique<PinnedLease>(ctx->pinned, buf_bytes));
    }
    std::vector<hipEvent_t> done_ev(static_cast<size_t>(nbuf), nullptr), copied_ev(static_cast<size_t>(nbuf), nullptr);
    auto ev_free = [&]() {
      for (int b = 0; b < nbuf; ++b)
      {
        if (done_ev[b])
          (void)hipEventDestroy(done_ev[b]);
        if (copied_ev[b])
          (void)hipEventDestroy(copied_ev[b]);
      }
    };
    struct Band
    {
      Set rows, prefix;
      DeviceBuffer sel_rows, sel_prefix;
      std::unique_ptr<sbamd_plan> cross, self;
    };
    std::vector<std::unique_ptr<Band>> band(static_cast<size_t>(nbands));
    const int rb = record_bytes(rs.tt_f64);
    uint64_t all_bytes = 0;
    {
      // records of the PCFs of original index below r: a running sum
      std::vector<int64_t> mrec(static_cast<size_t>(n));
      for (int64_t q = 0; q < n; ++q)
        mrec[static_cast<size_t>(rs.h_perm[q])] = rs.h_roff[q + 1] - rs.h_roff[q];
      uint64_t below = 0;
      int64_t r = 0;
      for (int64_t k = 0; k < nbands; ++k)
      {
        const uint64_t pre = below;
        for (; r < bounds[k + 1]; ++r)
          below += static_cast<uint64_t>(mrec[static_cast<size_t>(r)]);
        all_bytes += (below - pre + (cs ? 0 : pre)) * static_cast<uint64_t>(rb);
      }
    }
    // Bands whose records all fit in a quarter of the free memory are kept
    // until the job ends (freeing device memory between launches waits for
    // the device); otherwise each is freed once its output is copied out.
    size_t free_b = 0, total_b = 0;
    SBAMD_HIP(hipMemGetInfo(&free_b, &total_b));
    const bool keep = all_bytes <= static_cast<uint64_t>(free_b) / 4;
    auto build = [&](int64_t k) {
      const int64_t r0 = bounds[k], r1 = bounds[k + 1];
      const int64_t base = -static_cast<int64_t>(cum(r0));
      auto bd = std::make_unique<Band>();
      std::vector<int64_t> sel, pre;
      for (int64_t q = 0; q < n; ++q)
      {
        const int64_t i = rs.h_perm[q];
        if (i >= r0 && i < r1)
          sel.push_back(q);
        else if (!cs && i < r0)
          pre.push_back(q);
      }
      make_subset(rs, sel, bd->rows, bd->sel_rows, st);
      if (cs)
        bd->cross = new_plan(ctx, &bd->rows, cs, op, p, SBAMD_OUT_DENSE, &sh, base);
      else
      {
        if (!pre.empty())
        {
          make_subset(rs, pre, bd->prefix, bd->sel_prefix, st);
          bd->cross = new_plan(ctx, &bd->rows, &bd->prefix, op, p, out_kind, &sh, base);
        }
        bd->self = new_plan(ctx, &bd->rows, nullptr, op, p, out_kind, &sh, base);
      }
      band[static_cast<size_t>(k)] = std::move(bd);
    };
    // band k used buffer k % nbuf: wait for its download and copy it out
    auto retire = [&](int64_t k) {
      if (k < 0)
        return;
      const int b = static_cast<int>(k % nbuf);
      SBAMD_HIP(hipEventSynchronize(copied_ev[b]));
      const uint64_t e0 = cum(bounds[k]), e1 = cum(bounds[k + 1]);
      parallel_copy(static_cast<unsigned char*>(out_host) + e0 * eb, hbuf[b]->ptr, static_cast<size_t>(e1 - e0) * eb);
      jp->progress += e1 - e0;
      if (!keep)
        band[static_cast<size_t>(k)].reset();
    };
    // Band k's download, issued once its kernels have finished (the host
    // waits while the device runs band k + 1): with a cross-stream event
    // wait instead, hipMemcpyAsync blocked the host on some HIP runtimes.
    auto download = [&](int64_t k) {
      const int b = static_cast<int>(k % nbuf);
      SBAMD_HIP(hipEventSynchronize(done_ev[b]));
      SBAMD_HIP(hipMemcpyAsync(hbuf[b]->ptr, dbuf[b].ptr, static_cast<size_t>(cum(bounds[k + 1]) - cum(bounds[k])) * eb,
                               hipMemcpyDeviceToHost, cst));
      SBAMD_HIP(hipEventRecord(copied_ev[b], cst));
    };
    // band plans are made while the device runs the bands before them
    AsyncUploads uploads(ctx, st);
    try
    {
      for (int b = 0; b < nbuf; ++b)
      {
        SBAMD_HIP(hipEventCreateWithFlags(&done_ev[b], hipEventDisableTiming));
        SBAMD_HIP(hipEventCreateWithFlags(&copied_ev[b], hipEventDisableTiming));
      }
      for (int64_t k = 0; k < nbands; ++k)
      {
        const int b = static_cast<int>(k % nbuf);
        if (jp->cancel.load())
          fail(SBAMD_ERR_CANCELLED, "cancelled");
        build(k);
        retire(k - nbuf);
        Band& bd = *band[static_cast<size_t>(k)];
        if (bd.cross)
          execute_range(*bd.cross, dbuf[b].ptr, out_dtype, nullptr, 0, work_units(*bd.cross), st);
        if (bd.self)
          execute_range(*bd.self, dbuf[b].ptr, out_dtype, nullptr, 0, work_units(*bd.self), st);
        SBAMD_HIP(hipEventRecord(done_ev[b], st));
        if (k > 0)
          download(k - 1);
      }
      download(nbands - 1);
      for (int64_t k = std::max<int64_t>(0, nbands - nbuf); k < nbands; ++k)
        retire(k);
      SBAMD_HIP(hipStreamSynchronize(st));
    }
    catch (...)
    {
      (void)hipStreamSynchronize(st);
      (void)hipStreamSynchronize(cst);
      ev_free();
      throw;
    }
    ev_free();
  }

  // A job's set from a view: small contiguous host sets through the host
  // packer, the rest through the GPU packer (which also checks the points
  // and reports the first offending PCF with the host rules' message).
  void make_set_view(Ctx* ctx, const sbamd_pcf_view& in, int32_t precision, Set& s)
  {
    const int64_t es = in.dtype == SBAMD_F64 ? 8 : 4;
    if (in.location == SBAMD_HOST && in.t_stride == es && in.v_stride == es)
    {
      sbamd_pcf_set set;
      set.n = in.n;
      set.offsets = in.offsets;
      set.t = in.t;
      set.v = in.v;
      set.dtype = in.dtype;
      make_set_auto(ctx, set, precision, s);
      return;
    }
    make_set_gpu(ctx, in, precision, s);
  }

  sbamd_pcf_view view_of(const sbamd_pcf_set& in)
  {
    sbamd_pcf_view v;
    v.n = in.n;
    v.offsets = in.offsets;
    v.t = in.t;
    v.v = in.v;
    v.t_stride = v.v_stride = in.dtype == SBAMD_F64 ? 8 : 4;
    v.dtype = in.dtype;
    v.location = SBAMD_HOST;
    return v;
  }

  void run_job(Ctx* ctx, sbamd_job* jp, const sbamd_pcf_view& rows, const sbamd_pcf_view* cols, int32_t op,
               double p, int32_t out_kind, int32_t precision, void* out_host, int32_t out_dtype)
  {
    Set rs, cs;
    make_set_view(ctx, rows, precision, rs);
    if (cols)
      make_set_view(ctx, *cols, precision, cs);
    const size_t eb = out_dtype == SBAMD_F64 ? 8 : 4;
    const uint64_t nr = static_cast<uint64_t>(rs.n);
    const uint64_t out_count = cols ? nr * static_cast<uint64_t>(cs.n)
                                    : (out_kind == SBAMD_OUT_SYMMAT ? nr * (nr + 1) / 2 : (nr > 1 ? nr * (nr - 1) / 2 : 0));
    const uint64_t out_bytes = out_count * eb;
    uint64_t budget = ctx->out_budget;
    if (budget == 0)
    {
      size_t free_b = 0, total_b = 0;
      SBAMD_HIP(hipMemGetInfo(&free_b, &total_b));
      budget = static_cast<uint64_t>(free_b) / 2;
    }
    if (out_bytes <= JOB_SINGLE_MAX_BYTES && out_bytes <= budget)
    {
      job_single(ctx, jp, rs, cols ? &cs : nullptr, op, p, out_kind, out_host, out_dtype);
      return;
    }
    // JOB_BUFFERS band buffers within the budget
    const uint64_t chunk = std::max<uint64_t>(budget / (JOB_BUFFERS * eb), 1);
    job_bands(ctx, jp, rs, cols ? &cs : nullptr, op, p, out_kind, out_host, out_dtype, chunk);
  }
}

namespace sbamd
{
  namespace
  {
    // The job thread of sbamd_pairwise_submit(_view); the views' memory stays
    // the caller's until the job ends.
    void submit_job(sbamd_ctx* ctx, const sbamd_pcf_view& rows, const sbamd_pcf_view* cols, int32_t op, double p,
                    int32_t out_kind, int32_t precision, void* out_host, int32_t out_dtype, sbamd_job** out)
    {
      need(ctx, "ctx");
      need(out, "out");
      *out = nullptr;
      check_out_dtype(out_dtype);
      if (op == SBAMD_OP_LP && !(p >= 1.0))
        fail(SBAMD_ERR_INVALID, "p must be >= 1.");
      (void)kernel_op(op, p);
      if (rows.n < 0 || (cols && cols->n < 0))
        fail(SBAMD_ERR_INVALID, "PCF set size must be non-negative");
      if (rows.n > 0 || (cols && cols->n > 0))
        need(out_host, "out_host");
      auto job = std::make_unique<sbamd_job>();
      sbamd_job* jp = job.get();
      const sbamd_pcf_view rcopy = rows;
      const bool has_cols = cols != nullptr;
      const sbamd_pcf_view ccopy = cols ? *cols : sbamd_pcf_view{};
      uint64_t total = 0;
      if (has_cols)
        total = static_cast<uint64_t>(rows.n) * static_cast<uint64_t>(cols->n);
      else if (out_kind == SBAMD_OUT_SYMMAT)
        total = static_cast<uint64_t>(rows.n) * (rows.n + 1) / 2;
      else
        total = rows.n > 1 ? tri(static_cast<uint64_t>(rows.n)) : 0;
      jp->total = total;
      jp->worker = std::thread([=]() {
        int status = SBAMD_OK;
        std::string err;
        try
        {
          SBAMD_HIP(hipSetDevice(ctx->device));
          if (total > 0)
            run_job(ctx, jp, rcopy, has_cols ? &ccopy : nullptr, op, p, out_kind, precision, out_host,
                    out_dtype);
        }
        catch (const Error& e)
        {
          status = e.code;
          err = e.msg;
        }
        catch (const std::exception& e)
        {
          status = SBAMD_ERR_RUNTIME;
          err = e.what();
        }
        std::lock_guard<std::mutex> lk(jp->mu);
        jp->status = status;
        jp->error = err;
        jp->done = true;
        jp->cv.notify_all();
      });
      *out = job.release();
    }
  }
}

extern "C"
{
  const char* sbamd_last_error(void) { return g_last_error.c_str(); }

  const char* sbamd_version(void) { return "stablebear_amd 0.1.0 (gfx950)"; }

  int sbamd_device_count(int* count)
  {
    return guarded([&] {
      need(count, "count");
      int n = 0;
      hipError_t e = hipGetDeviceCount(&n);
      if (e != hipSuccess)
        n = 0;
      *count = n;
    });
  }

  int sbamd_ctx_create(int device, sbamd_ctx** out)
  {
    return guarded([&] {
      need(out, "out");
      *out = nullptr;
      int n = 0;
      if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        fail(SBAMD_ERR_NODEVICE, "no HIP device is available");
      if (device < 0 || device >= n)
        fail(SBAMD_ERR_INVALID, "device ordinal out of range");
      auto ctx = std::make_unique<sbamd_ctx>();
      ctx->device = device;
      SBAMD_HIP(hipSetDevice(device));
      hipDeviceProp_t prop;
      SBAMD_HIP(hipGetDeviceProperties(&prop, device));
      ctx->cu_count = prop.multiProcessorCount;
      ctx->lds_per_block = prop.sharedMemPerBlock;
      ctx->arch = prop.gcnArchName;
      {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        auto it = scratch_addr().find(device);
        if (it == scratch_addr().end())
        {
          void* ptr = nullptr;
          SBAMD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
          if (!ptr)
            fail(SBAMD_ERR_RUNTIME, "hipGetSymbolAddress returned no address for the working memory");
          it = scratch_addr().emplace(device, ptr).first;
        }
        ctx->scratch = it->second;
      }
      SBAMD_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
      *out = ctx.release();
    });
  }

  int sbamd_ctx_destroy(sbamd_ctx* ctx)
  {
    return guarded([&] {
      if (!ctx)
        return;
      use_device(ctx);
      if (ctx->stream)
      {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
      }
      delete ctx;
    });
  }

  int sbamd_ctx_synchronize(sbamd_ctx* ctx)
  {
    return guarded([&] {
      need(ctx, "ctx");
      use_device(ctx);
      SBAMD_HIP(hipStreamSynchronize(ctx->stream));
    });
  }

  int sbamd_ctx_set_output_budget(sbamd_ctx* ctx, uint64_t bytes)
  {
    return guarded([&] {
      need(ctx, "ctx");
      ctx->out_budget = bytes;
    });
  }

  int sbamd_ctx_set_tile_shape(sbamd_ctx* ctx, const sbamd_tile_shape* shape)
  {
    return guarded([&] {
      need(ctx, "ctx");
      need(shape, "shape");
      const int R = shape->rows ? shape->rows : 16, C = shape->cols ? shape->cols : 16;
      if (!tile_shape_ok(R, C, shape->cap ? shape->cap : 48, shape->lane_pairs ? shape->lane_pairs : 1))
        fail(SBAMD_ERR_INVALID, "tile shape out of range (rows + cols <= 64, lane_pairs 1 or 2, "
                                "at most 512 lanes, 2 <= cap <= 4096)");
      ctx->job_R = shape->rows;
      ctx->job_C = shape->cols;
      ctx->job_cap = shape->cap;
      ctx->job_ppl = shape->lane_pairs;
    });
  }

  int sbamd_ctx_stream(sbamd_ctx* ctx, void** stream)
  {
    return guarded([&] {
      need(ctx, "ctx");
      need(stream, "stream");
      *stream = static_cast<void*>(ctx->stream);
    });
  }

  int sbamd_validate_set(const sbamd_pcf_set* set)
  {
    return guarded([&] {
      need(set, "set");
      validate_set(*set);
    });
  }

  int sbamd_set_create(sbamd_ctx* ctx, const sbamd_pcf_set* set, int32_t precision, sbamd_set** out)
  {
    return guarded([&] {
      need(ctx, "ctx");
      need(set, "set");
      need(out, "out");
      *out = nullptr;
      auto s = std::make_unique<sbamd_set>();
      make_set(ctx, *set, precision, *s);
      *out = s.release();
    });
  }

  int sbamd_set_create_gpu(sbamd_ctx* ctx, const sbamd_pcf_view* view, int32_t precision, sbamd_set** out)
  {
    return guarded([&] {
      need(ctx, "ctx");
      need(view, "view");
      need(out, "out");
      *out = nullptr;
      auto s = std::make_unique<sbamd_set>();
      make_set_gpu(ctx, *view, precision, *s);
      *out = s.release();
    });
  }

  int sbamd_set_create_image(sbamd_ctx* ctx, const void* buf, uint64_t len, int32_t precision, sbamd_set** out)
  {
    return guarded([&] {
      need(ctx, "ctx");
      need(out, "out");
      if (!buf && len)
        fail(SBAMD_ERR_INVALID, "buf is NULL");
      *out = nullptr;
      auto s = std::make_unique<sbamd_set>();
      set_from_image(ctx, buf, len, precision, *s);
      *out = s.release();
    });
  }

  int sbamd_rows_reduce(sbamd_ctx* ctx, int32_t out_kind, int64_t n_rows, int64_t n_cols, const void* d_device,
                        int32_t dtype, double* sum_device, double* min_device, double* max_device, void* stream)
  {
    return guarded([&] {
      need(ctx, "ctx");
      RowsParams P = rows_params(out_kind, n_rows, n_cols, d_device, dtype);
      P.sum = sum_device;
      P.min = min_device;
      P.max = max_device;
      use_device(ctx);
      launch_rows_reduce(P, pick_stream(ctx, stream));
    });
  }

  int sbamd_rows_knn(sbamd_ctx* ctx, int32_t out_kind, int64_t n_rows, int64_t n_cols, const void* d_device,
                     int32_t dtype, int32_t k, void* values_device, int64_t* index_device, void* stream)
  {
    return guarded([&] {
      need(ctx, "ctx");
      if (k < 1 || k > 32)
        fail(SBAMD_ERR_INVALID, "k must be in [1, 32]");
      RowsParams P = rows_params(out_kind, n_rows, n_cols, d_device, dtype);
      if (P.n_rows > 0)
      {
        need(values_device, "values_device");
        need(index_device, "index_device");
      }
      P.k = k;
      P.knn_v = values_device;
      P.knn_i = index_device;
      use_device(ctx);
      launch_rows_knn(P, pick_stream(ctx, stream));
    });
  }

  int sbamd_io_parse(const void* buf, uint64_t len, sbamd_io_info* info)
  {
    return guarded([&] {
      need(info, "info");
      io_parse(buf, len, *info);
    });
  }

  int sbamd_io_index(const void* buf, uint64_t len, const sbamd_io_info* info, uint64_t* counts,
                     uint64_t* starts, uint64_t* end)
  {
    return guarded([&] {
      need(info, "info");
      const uint64_t e = io_index(buf, len, *info, counts, starts);
      if (end)
        *end = e;
    });
  }

  int sbamd_io_header(uint32_t kind, int32_t base_format, int32_t sub_format, int32_t ndim,
                      const uint64_t* shape, void* out, uint64_t cap, uint64_t* size)
  {
    return guarded([&] {
      const uint64_t n = io_header(kind, base_format, sub_format, ndim, shape, out, cap);
      if (size)
        *size = n;
    });
  }

  int sbamd_io_write_pcfs(int64_t n, const int64_t* offsets, const void* t, const void* v, int32_t dtype,
                          void* out, uint64_t cap, uint64_t* size)
  {
    return guarded([&] {
      const uint64_t b = io_write_pcfs(n, offsets, t, v, dtype, out, cap);
      if (size)
        *size = b;
    });
  }

  int sbamd_set_destroy(sbamd_set* set)
  {
    return guarded([&] {
      if (!set)
        return;
      use_device(set->ctx);
      delete set;
    });
  }

  int sbamd_set_info(const sbamd_set* set, int64_t* n, int64_t* records, int64_t* points)
  {
    return guarded([&] {
      need(set, "set");
      if (n)
        *n = set->n;
      if (records)
        *records = set->records;
      if (points)
        *points = set->points;
    });
  }

  int sbamd_set_export(const sbamd_set* set, void* rec, int64_t* roff, int64_t* perm, int64_t* src,
                       int32_t* rec_bytes)
  {
    return guarded([&] {
      need(set, "set");
      use_device(set->ctx);
      SBAMD_HIP(hipDeviceSynchronize());
      auto get = [](void* dst, const DeviceBuffer& b) {
        if (dst && b.bytes)
          SBAMD_HIP(hipMemcpy(dst, b.ptr, b.bytes, hipMemcpyDeviceToHost));
      };
      get(rec, set->rec);
      get(roff, set->roff);
      get(perm, set->perm);
      get(src, set->src);
      if (rec_bytes)
        *rec_bytes = record_bytes(set->tt_f64);
    });
  }

  int sbamd_set_update_values(sbamd_set* set, const void* v_device, void* stream)
  {
    return guarded([&] {
      need(set, "set");
      need(v_device, "v_device");
      use_device(set->ctx);
      if (set->records > 0 && set->points == 0)
        return; // only zero PCFs: nothing to refresh
      if (set->has_empty)
        fail(SBAMD_ERR_INVALID, "value updates need sets without empty PCFs");
      launch_update_values(v_device, set->src.as<int64_t>(), set->records, set->rec.ptr,
                           record_bytes(set->tt_f64), set->in_dtype == SBAMD_F64, set->tv_f64,
                           pick_stream(set->ctx, stream));
      ++set->vgen;
    });
  }

  int sbamd_set_lp_norm(sbamd_set* set, double p, void* out_device, int32_t out_dtype, void* stream)
  {
    return guarded([&] {
      need(set, "set");
      check_out_dtype(out_dtype);
      if (!(p >= 1.0))
        fail(SBAMD_ERR_INVALID, "p must be >= 1.");
      if (set->n == 0)
        return;
      need(out_device, "out_device");
      use_device(set->ctx);
      NormParams prm{};
      prm.roff = set->roff.as<int64_t>();
      prm.rec = set->rec.ptr;
      prm.perm = set->perm.as<int64_t>();
      prm.n = set->n;
      prm.p = p;
      prm.out = out_device;
      prm.out_f64 = out_dtype == SBAMD_F64;
      launch_norm(prm, set->tt_f64, set->tv_f64, pick_stream(set->ctx, stream));
    });
  }

  int sbamd_plan_create(sbamd_ctx* ctx, const sbamd_set* rows, const sbamd_set* cols, int32_t op,
                        double p, int32_t out_kind, const sbamd_tile_shape* shape, sbamd_plan** out)
  {
    return guarded([&] {
      need(ctx, "ctx");
      need(rows, "rows");
      need(out, "out");
      *out = nullptr;
      if (cols && out_kind != SBAMD_OUT_DENSE)
        fail(SBAMD_ERR_INVALID, "cross plans (cols != NULL) need SBAMD_OUT_DENSE");
      *out = new_plan(ctx, rows, cols, op, p, out_kind, shape, 0).release();
    });
  }

  int sbamd_plan_create_batched(sbamd_ctx* ctx, int32_t k, const sbamd_set* const* sets,
                                int32_t op, double p, int32_t out_kind,
                                const sbamd_tile_shape* shape, sbamd_plan** out)
  {
    return guarded([&] {
      need(ctx, "ctx");
      need(sets, "sets");
      need(out, "out");
      *out = nullptr;
      if (k < 1 || k > 65535)
        fail(SBAMD_ERR_INVALID, "batch size must be in [1, 65535]");
      if (out_kind != SBAMD_OUT_DISTMAT && out_kind != SBAMD_OUT_SYMMAT)
        fail(SBAMD_ERR_INVALID, "batched plans are pdist / kernel (DISTMAT or SYMMAT)");
      if (op == SBAMD_OP_LP && !(p >= 1.0))
        fail(SBAMD_ERR_INVALID, "p must be >= 1.");
      (void)kernel_op(op, p);
      auto pl = std::make_unique<sbamd_plan>();
      pl->ctx = ctx;
      pl->batched = true;
      pl->persistent = true;
      pl->op = op;
      pl->p = op == SBAMD_OP_LP ? p : 1.0;
      pl->out_kind = out_kind;
      need(sets[0], "sets[0]");
      pl->precision = sets[0]->precision;
      pl->tt_f64 = sets[0]->tt_f64;
      pl->tv_f64 = sets[0]->tv_f64;
      int64_t pcf_base = 0, rec_base = 0;
      uint64_t out_base = 0;
      std::vector<int64_t> roff_all{0};
      std::vector<int64_t> perm_all;
      for (int32_t b = 0; b < k; ++b)
      {
        const Set* s = sets[b];
        need(s, "sets[b]");
        if (s->precision != pl->precision || s->tt_f64 != pl->tt_f64 || s->tv_f64 != pl->tv_f64)
          fail(SBAMD_ERR_INVALID, "all sets of a batch must share dtype and precision");
        pl->sets.push_back(s);
        Problem pr{pcf_base, pcf_base, static_cast<int64_t>(out_base), s->n};
        pl->h_probs.push_back(pr);
        const uint64_t n = static_cast<uint64_t>(s->n);
        const uint64_t M = static_cast<uint64_t>(s->records);
        uint64_t pairs = out_kind == SBAMD_OUT_DISTMAT ? (n > 1 ? tri(n) : 0) : n * (n + 1) / 2;
        pl->pairs += pairs;
        if (out_kind == SBAMD_OUT_DISTMAT)
          pl->segments += n > 1 ? (n - 1) * M - pairs : 0;
        else
          pl->segments += n > 0 ? n * M - (n > 1 ? tri(n) : 0) : 0;
        out_base += pairs;
        for (int64_t q = 0; q < s->n; ++q)
        {
          roff_all.push_back(rec_base + s->h_roff[q + 1]);
          perm_all.push_back(s->h_perm[q]);
        }
        pcf_base += s->n;
        rec_base += s->records;
      }
      pl->out_count = out_base;
      use_device(ctx);
      const size_t rb = static_cast<size_t>(record_bytes(pl->tt_f64));
      pl->b_rec = DeviceBuffer(static_cast<size_t>(rec_base) * rb);
      int64_t at = 0;
      for (int32_t b = 0; b < k; ++b)
      {
        const Set* s = sets[b];
        pl->b_at.push_back(at);
        pl->b_vgen.push_back(s->vgen);
        if (s->records == 0)
          continue;
        SBAMD_HIP(hipMemcpy(static_cast<char*>(pl->b_rec.ptr) + at * rb, s->rec.ptr,
                            static_cast<size_t>(s->records) * rb, hipMemcpyDeviceToDevice));
        at += s->records;
      }
      upload(pl->b_roff, roff_all);
      upload(pl->b_perm, perm_all);
      pl->input_bytes = pl->b_rec.bytes + pl->b_roff.bytes;
      set_default_shape(*pl);
      if (shape)
      {
        if (shape->rows)
          pl->R = shape->rows;
        if (shape->cols)
          pl->C = shape->cols;
        else if (shape->rows)
          pl->C = pl->R; // only a side given: a square tile
        if (shape->cap)
          pl->cap = shape->cap;
        if (shape->lane_pairs)
          pl->ppl = shape->lane_pairs;
      }
      check_shape(pl->R, pl->C, pl->cap, pl->ppl);
      pl->rebuild_tiles();
      *out = pl.release();
    });
  }

  int sbamd_plan_destroy(sbamd_plan* plan)
  {
    return guarded([&] {
      if (!plan)
        return;
      forget_plan(plan);
      use_device(plan->ctx);
      delete plan;
    });
  }

  int sbamd_plan_info(const sbamd_plan* plan, sbamd_plan_stats* info)
  {
    return guarded([&] {
      need(plan, "plan");
      need(info, "info");
      info->pairs = plan->pairs;
      info->segments = plan->segments;
      info->out_count = plan->out_count;
      info->tiles = plan->h_tiles.size();
      info->input_bytes = plan->input_bytes;
      // with per-class shapes: the class with the most tiles
      int kb = 0;
      for (int k = 1; k < NCLASS; ++k)
        if (plan->class_begin[k + 1] - plan->class_begin[k] > plan->class_begin[kb + 1] - plan->class_begin[kb])
          kb = k;
      const Shape sh = plan->shape_of(kb);
      info->rows = sh.R;
      info->cols = sh.C;
      info->cap = sh.cap;
      info->problems = static_cast<int32_t>(plan->h_probs.size());
      info->long_pairs = plan->use_long ? 1 : 0;
      info->lane_pairs = sh.ppl;
    });
  }

  int sbamd_plan_set_long_pairs(sbamd_plan* plan, int32_t mode)
  {
    return guarded([&] {
      need(plan, "plan");
      if (mode < 0 || mode > 2)
        fail(SBAMD_ERR_INVALID, "long-pair mode must be 0 (off), 1 (on) or 2 (auto)");
      if (mode == 1 && plan->batched)
        fail(SBAMD_ERR_INVALID, "batched plans use the tile kernel");
      plan->long_mode = mode;
      decide_long(*plan);
    });
  }

  int sbamd_plan_out_offset(const sbamd_plan* plan, int32_t problem, uint64_t* offset)
  {
    return guarded([&] {
      need(plan, "plan");
      need(offset, "offset");
      if (problem < 0 || problem >= static_cast<int32_t>(plan->h_probs.size()))
        fail(SBAMD_ERR_INVALID, "problem index out of range");
      *offset = static_cast<uint64_t>(plan->h_probs[problem].out_base);
    });
  }

  int sbamd_plan_set_fold_threshold(sbamd_plan* plan, int64_t segments)
  {
    return guarded([&] {
      need(plan, "plan");
      if (segments < 1)
        fail(SBAMD_ERR_INVALID, "fold threshold must be >= 1");
      plan->fold_threshold = segments;
      plan->rebuild_tiles();
    });
  }

  int sbamd_plan_set_persistent(sbamd_plan* plan, int32_t on)
  {
    return guarded([&] {
      need(plan, "plan");
      plan->persistent = on != 0;
      plan->bump();
    });
  }

  int sbamd_plan_set_shape(sbamd_plan* plan, const sbamd_tile_shape* shape)
  {
    return guarded([&] {
      need(plan, "plan");
      need(shape, "shape");
      int R = shape->rows ? shape->rows : plan->R;
      int C = shape->cols ? shape->cols : (shape->rows && !plan->cdist ? R : plan->C);
      int cap = shape->cap ? shape->cap : plan->cap;
      int ppl = shape->lane_pairs ? shape->lane_pairs : plan->ppl;
      check_shape(R, C, cap, ppl);
      plan->R = R;
      plan->C = C;
      plan->cap = cap;
      plan->ppl = ppl;
      plan->per_class = false;
      plan->rebuild_tiles();
    });
  }

  int sbamd_plan_execute(sbamd_plan* plan, void* out_device, int32_t out_dtype, void* row_sum_device,
                         void* stream)
  {
    return guarded([&] {
      need(plan, "plan");
      check_out_dtype(out_dtype);
      if (plan->out_count > 0)
        need(out_device, "out_device");
      if (row_sum_device && !plan->cdist)
        fail(SBAMD_ERR_INVALID, "row sums are available for cross (dense) plans only");
      use_device(plan->ctx);
      hipStream_t st = pick_stream(plan->ctx, stream);
      if (row_sum_device)
        launch_fill_zero(row_sum_device, static_cast<size_t>(plan->sets[0]->n) * sizeof(double), st);
      if (plan->batched)
        refresh_batch(*plan, st);
      execute_range(*plan, out_device, out_dtype, static_cast<double*>(row_sum_device), 0,
                    work_units(*plan), st);
    });
  }

  int sbamd_plan_execute_to_host(sbamd_plan* plan, void* out_host, int32_t out_dtype)
  {
    return guarded([&] {
      need(plan, "plan");
      check_out_dtype(out_dtype);
      if (plan->out_count == 0)
        return;
      need(out_host, "out_host");
      use_device(plan->ctx);
      const size_t eb = out_dtype == SBAMD_F64 ? 8 : 4;
      DeviceBuffer dout(static_cast<size_t>(plan->out_count) * eb);
      if (plan->batched)
        refresh_batch(*plan, plan->ctx->stream);
      execute_range(*plan, dout.ptr, out_dtype, nullptr, 0, work_units(*plan),
                    plan->ctx->stream);
      SBAMD_HIP(hipStreamSynchronize(plan->ctx->stream));
      SBAMD_HIP(hipMemcpy(out_host, dout.ptr, static_cast<size_t>(plan->out_count) * eb,
                          hipMemcpyDeviceToHost));
    });
  }

  int sbamd_plain_create(sbamd_ctx* ctx, const sbamd_pcf_set* set, sbamd_plain** out)
  {
    return guarded([&] {
      need(ctx, "ctx");
      need(set, "set");
      need(out, "out");
      *out = nullptr;
      if (set->dtype != SBAMD_F64)
        fail(SBAMD_ERR_INVALID, "the plain FP64 kernel takes float64 PCF sets");
      validate_set(*set);
      auto pl = std::make_unique<sbamd_plain>();
      pl->ctx = ctx;
      pl->n = set->n;
      use_device(ctx);
      const int64_t M = set->n > 0 ? set->offsets[set->n] : 0;
      std::vector<int64_t> off(set->offsets, set->offsets + set->n + 1);
      std::vector<double> t(static_cast<const double*>(set->t), static_cast<const double*>(set->t) + M);
      std::vector<double> v(static_cast<const double*>(set->v), static_cast<const double*>(set->v) + M);
      upload(pl->off, off);
      upload(pl->t, t);
      upload(pl->v, v);
      *out = pl.release();
    });
  }

  int sbamd_plain_execute(sbamd_plain* plain, int32_t op, double p, int32_t out_kind, void* out_device,
                          void* stream)
  {
    return guarded([&] {
      need(plain, "plain");
      if (op == SBAMD_OP_LP && !(p >= 1.0))
        fail(SBAMD_ERR_INVALID, "p must be >= 1.");
      if (out_kind != SBAMD_OUT_DISTMAT && out_kind != SBAMD_OUT_SYMMAT)
        fail(SBAMD_ERR_INVALID, "the plain kernel writes DISTMAT or SYMMAT");
      if (plain->n > 1 || (plain->n == 1 && out_kind == SBAMD_OUT_SYMMAT))
        need(out_device, "out_device");
      use_device(plain->ctx);
      launch_plain(plain->off.as<int64_t>(), plain->t.as<double>(), plain->v.as<double>(), plain->n,
                   kernel_op(op, p), op == SBAMD_OP_LP ? p : 1.0, out_kind,
                   static_cast<double*>(out_device), pick_stream(plain->ctx, stream));
    });
  }

  int sbamd_plain_destroy(sbamd_plain* plain)
  {
    return guarded([&] {
      if (!plain)
        return;
      use_device(plain->ctx);
      delete plain;
    });
  }

  int sbamd_plan_backward(sbamd_plan* plan, const void* grad_out_device, const void* fwd_out_device,
                          int32_t out_dtype, void* grad_v_device, void* grad_t_device,
                          int32_t grad_dtype, void* stream)
  {
    return guarded([&] {
      const int kop = backward_checks(plan, grad_out_device, fwd_out_device, out_dtype, grad_v_device,
                                      grad_t_device, grad_dtype);
      use_device(plan->ctx);
      const Set* s = plan->sets[0];
      const size_t M = static_cast<size_t>(s->records);
      const size_t N = static_cast<size_t>(s->n);
      if (plan->bwd_q.bytes < M * 8)
      {
        plan->bwd_key0 = DeviceBuffer(M * 8);
        plan->bwd_val0 = DeviceBuffer(M * 8);
        plan->bwd_col0 = DeviceBuffer(M * 4);
        plan->bwd_src0 = DeviceBuffer(M * 8);
        for (int b = 0; b < 2; ++b)
        {
          plan->bwd_keys[b] = DeviceBuffer(M * 8);
          plan->bwd_ids[b] = DeviceBuffer(M * 4);
        }
        plan->bwd_events = DeviceBuffer(M * sizeof(BwdEvent));
        plan->bwd_events2 = DeviceBuffer((M + BWD_EVENT_PAD) * sizeof(BwdEvent));
        plan->bwd_q = DeviceBuffer(M * 8);
        plan->bwd_a = DeviceBuffer(M * 8);
        plan->bwd_w = DeviceBuffer(N * 8);
        plan->bwd_ties = DeviceBuffer(64);
        plan->bwd_lists_ready = false;
      }
      BwdWork wk;
      wk.key0 = plan->bwd_key0.as<double>();
      wk.val0 = plan->bwd_val0.as<double>();
      wk.col0 = plan->bwd_col0.as<int32_t>();
      wk.src0 = plan->bwd_src0.as<int64_t>();
      for (int b = 0; b < 2; ++b)
      {
        wk.keys[b] = plan->bwd_keys[b].as<double>();
        wk.ids[b] = plan->bwd_ids[b].as<int32_t>();
      }
      wk.events = plan->bwd_events.as<BwdEvent>();
      wk.events2 = plan->bwd_events2.as<BwdEvent>();
      wk.q = plan->bwd_q.as<double>();
      wk.a = plan->bwd_a.as<double>();
      wk.w = plan->bwd_w.as<double>();
      wk.ties = plan->bwd_ties.as<int32_t>();
      run_backward(*plan, kop, wk, plan->bwd_lists_ready, grad_out_device, fwd_out_device, out_dtype,
                   grad_v_device, grad_t_device, grad_dtype, pick_stream(plan->ctx, stream));
      plan->bwd_lists_ready = true;
    });
  }

  int sbamd_plan_backward_noalloc(sbamd_plan* plan, const void* grad_out_device, const void* fwd_out_device,
                                  int32_t out_dtype, void* grad_v_device, void* grad_t_device,
                                  int32_t grad_dtype, void* stream)
  {
    return guarded([&] {
      const int kop = backward_checks(plan, grad_out_device, fwd_out_device, out_dtype, grad_v_device,
                                      grad_t_device, grad_dtype);
      const Set* s = plan->sets[0];
      if (s->records > NOALLOC_MAX_RECORDS || s->n > NOALLOC_MAX_PCFS)
        fail(SBAMD_ERR_INVALID, "sbamd_plan_backward_noalloc: the set has " + std::to_string(s->n) +
                                    " PCFs and " + std::to_string(s->records) +
                                    " records; its working memory holds at most " +
                                    std::to_string(NOALLOC_MAX_PCFS) + " PCFs and " +
                                    std::to_string(NOALLOC_MAX_RECORDS) + " records");
      use_device(plan->ctx);
      unsigned char* const base = static_cast<unsigned char*>(plan->ctx->scratch);
      if (!base)
        fail(SBAMD_ERR_RUNTIME, "no working memory address for this device");
      BwdWork wk;
      wk.key0 = reinterpret_cast<double*>(base + NOALLOC_KEY0);
      wk.val0 = reinterpret_cast<double*>(base + NOALLOC_VAL0);
      wk.src0 = reinterpret_cast<int64_t*>(base + NOALLOC_SRC0);
      wk.keys[0] = reinterpret_cast<double*>(base + NOALLOC_KEYS0);
      wk.keys[1] = reinterpret_cast<double*>(base + NOALLOC_KEYS1);
      wk.q = reinterpret_cast<double*>(base + NOALLOC_Q);
      wk.a = reinterpret_cast<double*>(base + NOALLOC_A);
      wk.events = reinterpret_cast<BwdEvent*>(base + NOALLOC_EVENTS);
      wk.events2 = reinterpret_cast<BwdEvent*>(base + NOALLOC_EVENTS2);
      wk.col0 = reinterpret_cast<int32_t*>(base + NOALLOC_COL0);
      wk.ids[0] = reinterpret_cast<int32_t*>(base + NOALLOC_IDS0);
      wk.ids[1] = reinterpret_cast<int32_t*>(base + NOALLOC_IDS1);
      wk.w = reinterpret_cast<double*>(base + NOALLOC_W);
      wk.ties = nullptr; // the lists are sorted on every call here, without the global order that rules ties out
      // the working array is shared by every plan on the device: sort every time
      run_backward(*plan, kop, wk, false, grad_out_device, fwd_out_device, out_dtype, grad_v_device,
                   grad_t_device, grad_dtype, pick_stream(plan->ctx, stream));
    });
  }

  int sbamd_graph_create(sbamd_plan* plan, void* out_device, int32_t out_dtype, sbamd_graph** out)
  {
    return guarded([&] {
      need(plan, "plan");
      need(out, "out");
      *out = nullptr;
      check_out_dtype(out_dtype);
      use_device(plan->ctx);
      auto g = std::make_unique<sbamd_graph>();
      g->ctx = plan->ctx;
      g->plan = plan;
      if (plan->batched)
        refresh_batch(*plan, plan->ctx->stream);
      g->gen = plan->gen;
      hipStream_t cs;
      SBAMD_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
      hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess)
      {
        try
        {
          execute_range(*plan, out_device, out_dtype, nullptr, 0,
                        work_units(*plan), cs);
        }
        catch (...)
        {
          hipGraph_t tmp = nullptr;
          (void)hipStreamEndCapture(cs, &tmp);
          if (tmp)
            (void)hipGraphDestroy(tmp);
          (void)hipStreamDestroy(cs);
          throw;
        }
        e = hipStreamEndCapture(cs, &g->graph);
      }
      (void)hipStreamDestroy(cs);
      SBAMD_HIP(e);
      SBAMD_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
      *out = g.release();
    });
  }

  int sbamd_graph_launch(sbamd_graph* graph, void* stream)
  {
    return guarded([&] {
      need(graph, "graph");
      if (!plan_current(graph->plan, graph->gen))
        fail(SBAMD_ERR_INVALID, "graph is stale: its plan was reshaped, retuned or destroyed "
                                "after capture; capture a new graph");
      use_device(graph->ctx);
      hipStream_t st = pick_stream(graph->ctx, stream);
      if (graph->plan->batched)
        refresh_batch(*graph->plan, st);
      SBAMD_HIP(hipGraphLaunch(graph->exec, st));
    });
  }

  int sbamd_graph_destroy(sbamd_graph* graph)
  {
    return guarded([&] {
      if (!graph)
        return;
      use_device(graph->ctx);
      if (graph->exec)
        (void)hipGraphExecDestroy(graph->exec);
      if (graph->graph)
        (void)hipGraphDestroy(graph->graph);
      delete graph;
    });
  }

  int sbamd_autotune(sbamd_plan* plan, void* out_device, int32_t out_dtype, int32_t reps,
                     sbamd_tile_shape* best, double* best_ms)
  {
    return guarded([&] {
      need(plan, "plan");
      check_out_dtype(out_dtype);
      if (plan->out_count > 0)
        need(out_device, "out_device");
      if (reps < 1)
        reps = 1;
      use_device(plan->ctx);
      hipStream_t st = plan->ctx->stream;
      if (plan->batched)
        refresh_batch(*plan, st);
      plan->use_long = false; // tile candidates first
      const std::vector<Shape> cands = tune_candidates(*plan);
      hipEvent_t e0, e1;
      SBAMD_HIP(hipEventCreate(&e0));
      SBAMD_HIP(hipEventCreate(&e1));
      auto egd = [&]() {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
      };
      // ms of tile range [t0, t1) (one warm-up launch, then reps)
      auto time_range = [&](int64_t t0, int64_t t1) {
        execute_range(*plan, out_device, out_dtype, nullptr, t0, t1, st);
        SBAMD_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r)
          execute_range(*plan, out_device, out_dtype, nullptr, t0, t1, st);
        SBAMD_HIP(hipEventRecord(e1, st));
        SBAMD_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        SBAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
        return static_cast<double>(ms) / reps;
      };
      double tiled_ms = 0.0;
      bool chose_long = false;
      try
      {
        // one shape per length class, each timed on its own class's launch
        if (!plan->per_class)
        {
          plan->per_class = true;
          for (int k = 0; k < NCLASS; ++k)
            plan->cls[k] = Shape{plan->R, plan->C, plan->cap, plan->ppl};
        }
        plan->rebuild_tiles();
        for (int k = 0; k < NCLASS; ++k)
        {
          if (plan->class_begin[k + 1] == plan->class_begin[k])
            continue;
          const std::string key = tune_key(*plan, k);
          Shape hit;
          if (tune_cache_get(key, hit))
          {
            plan->cls[k] = hit;
            plan->rebuild_tiles();
            continue;
          }
          double bestv = std::numeric_limits<double>::infinity();
          Shape bsh = plan->cls[k];
          for (const Shape& c : cands)
          {
            plan->cls[k] = c;
            plan->rebuild_tiles();
            const double ms = time_range(plan->class_begin[k], plan->class_begin[k + 1]);
            if (ms < bestv)
            {
              bestv = ms;
              bsh = c;
            }
          }
          plan->cls[k] = bsh;
          plan->rebuild_tiles();
          tune_cache_put(key, bsh);
        }
        tiled_ms = time_range(0, work_units(*plan));
        // Separate class launches each end with a tail of their longest
        // tiles, which one longest-first launch fills with shorter ones: one
        // shape for every class (the plan's own, or a class winner) in one
        // launch stays when it is faster.
        {
          const Shape tuned[NCLASS] = {plan->cls[0], plan->cls[1], plan->cls[2], plan->cls[3]};
          std::vector<Shape> singles{Shape{plan->R, plan->C, plan->cap, plan->ppl}};
          for (int k = 0; k < NCLASS; ++k)
          {
            bool seen = false;
            for (const Shape& x : singles)
              seen |= x.R == tuned[k].R && x.C == tuned[k].C && x.cap == tuned[k].cap && x.ppl == tuned[k].ppl;
            if (!seen && plan->class_begin[k + 1] > plan->class_begin[k])
              singles.push_back(tuned[k]);
          }
          Shape best_single{};
          double best_single_ms = std::numeric_limits<double>::infinity();
          plan->per_class = false;
          for (const Shape& x : singles)
          {
            plan->R = x.R;
            plan->C = x.C;
            plan->cap = x.cap;
            plan->ppl = x.ppl;
            plan->rebuild_tiles();
            const double ms = time_range(0, work_units(*plan));
            if (ms < best_single_ms)
            {
              best_single_ms = ms;
              best_single = x;
            }
          }
          plan->R = best_single.R;
          plan->C = best_single.C;
          plan->cap = best_single.cap;
          plan->ppl = best_single.ppl;
          if (best_single_ms <= tiled_ms)
          {
            tiled_ms = best_single_ms;
            plan->rebuild_tiles();
          }
          else
          {
            plan->per_class = true;
            plan->rebuild_tiles();
          }
        }
        // the long-pair kernel is a candidate too (FP32C plans, or plans
        // that asked for it)
        if (!plan->batched && plan->pairs > 0 &&
            (plan->precision == SBAMD_PREC_FP32C || plan->long_mode == 1))
        {
          plan->use_long = true;
          const double lms = time_range(0, work_units(*plan));
          chose_long = lms < tiled_ms;
          if (chose_long)
            tiled_ms = lms;
        }
      }
      catch (...)
      {
        egd();
        throw;
      }
      egd();
      plan->use_long = chose_long;
      plan->long_mode = chose_long ? 1 : 0;
      plan->bump();
      if (best)
      {
        // the shape of the class with the most tiles
        int kb = 0;
        for (int k = 1; k < NCLASS; ++k)
          if (plan->class_begin[k + 1] - plan->class_begin[k] > plan->class_begin[kb + 1] - plan->class_begin[kb])
            kb = k;
        const Shape sh = plan->shape_of(kb);
        best->rows = sh.R;
        best->cols = sh.C;
        best->cap = sh.cap;
        best->lane_pairs = sh.ppl;
      }
      if (best_ms)
        *best_ms = tiled_ms;
    });
  }

  int sbamd_plan_set_class_shapes(sbamd_plan* plan, const sbamd_tile_shape* shapes)
  {
    return guarded([&] {
      need(plan, "plan");
      need(shapes, "shapes");
      Shape sh[NCLASS];
      for (int k = 0; k < NCLASS; ++k)
      {
        const int R = shapes[k].rows ? shapes[k].rows : plan->R;
        sh[k] = Shape{R, shapes[k].cols ? shapes[k].cols : (shapes[k].rows && !plan->cdist ? R : plan->C),
                      shapes[k].cap ? shapes[k].cap : plan->cap,
                      shapes[k].lane_pairs ? shapes[k].lane_pairs : plan->ppl};
        check_shape(sh[k].R, sh[k].C, sh[k].cap, sh[k].ppl);
      }
      for (int k = 0; k < NCLASS; ++k)
        plan->cls[k] = sh[k];
      plan->per_class = true;
      plan->rebuild_tiles();
    });
  }

  int sbamd_autotune_cache_stats(uint64_t* hits, uint64_t* misses)
  {
    return guarded([&] {
      if (hits)
        *hits = g_tune_hits.load();
      if (misses)
        *misses = g_tune_misses.load();
    });
  }

  int sbamd_plan_class_shapes(const sbamd_plan* plan, sbamd_tile_shape* shapes, int64_t* tiles)
  {
    return guarded([&] {
      need(plan, "plan");
      for (int k = 0; k < NCLASS; ++k)
      {
        // the class's own shape once one was set or tuned (kept even when
        // the plan runs every class in one launch), else the plan's
        const Shape sh = plan->per_class || plan->cls[k].R ? plan->cls[k] : plan->shape_of(k);
        if (shapes)
        {
          shapes[k].rows = sh.R;
          shapes[k].cols = sh.C;
          shapes[k].cap = sh.cap;
          shapes[k].lane_pairs = sh.ppl;
        }
        if (tiles)
          tiles[k] = plan->class_begin[k + 1] - plan->class_begin[k];
      }
    });
  }

  int sbamd_pairwise_submit(sbamd_ctx* ctx, const sbamd_pcf_set* rows, const sbamd_pcf_set* cols,
                            int32_t op, double p, int32_t out_kind, int32_t precision, void* out_host,
                            int32_t out_dtype, sbamd_job** out)
  {
    return guarded([&] {
      need(rows, "rows");
      // Validate on the caller's thread so input errors surface at submit.
      validate_set(*rows);
      if (cols)
        validate_set(*cols);
      const sbamd_pcf_view rv = view_of(*rows);
      const sbamd_pcf_view cv = cols ? view_of(*cols) : sbamd_pcf_view{};
      submit_job(ctx, rv, cols ? &cv : nullptr, op, p, out_kind, precision, out_host, out_dtype, out);
    });
  }

  int sbamd_pairwise_submit_view(sbamd_ctx* ctx, const sbamd_pcf_view* rows, const sbamd_pcf_view* cols,
                                 int32_t op, double p, int32_t out_kind, int32_t precision, void* out_host,
                                 int32_t out_dtype, sbamd_job** out)
  {
    return guarded([&] {
      need(rows, "rows");
      // the offsets here; the points on the device, during the job
      for (const sbamd_pcf_view* v : {rows, cols})
        if (v)
        {
          if (v->location != SBAMD_HOST && v->location != SBAMD_DEVICE)
            fail(SBAMD_ERR_INVALID, "PCF view location must be SBAMD_HOST or SBAMD_DEVICE");
          if (v->location == SBAMD_HOST)
            check_set_offsets(v->n, v->offsets, v->dtype);
        }
      submit_job(ctx, *rows, cols, op, p, out_kind, precision, out_host, out_dtype, out);
    });
  }


  int sbamd_job_wait(sbamd_job* job, int64_t timeout_ms)
  {
    if (!job)
    {
      g_last_error = "job is NULL";
      return SBAMD_ERR_INVALID;
    }
    std::unique_lock<std::mutex> lk(job->mu);
    if (timeout_ms < 0)
      job->cv.wait(lk, [&] { return job->done; });
    else if (!job->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return job->done; }))
      return SBAMD_TIMEOUT;
    if (job->status != SBAMD_OK)
      g_last_error = job->error;
    return job->status;
  }

  int sbamd_job_progress(sbamd_job* job, uint64_t* done, uint64_t* total)
  {
    return guarded([&] {
      need(job, "job");
      if (done)
        *done = job->progress.load();
      if (total)
        *total = job->total.load();
    });
  }

  int sbamd_job_cancel(sbamd_job* job)
  {
    return guarded([&] {
      need(job, "job");
      job->cancel = true;
    });
  }

  int sbamd_job_destroy(sbamd_job* job)
  {
    return guarded([&] {
      if (!job)
        return;
      job->cancel = true;
      if (job->worker.joinable())
        job->worker.join();
      delete job;
    });
  }

  int sbamd_lp_distance(sbamd_ctx* ctx, const sbamd_pcf_set* f, const sbamd_pcf_set* g, double p,
                        double* out)
  {
    return guarded([&] {
      need(ctx, "ctx");
      need(f, "f");
      need(g, "g");
      need(out, "out");
      if (f->n != 1 || g->n != 1)
        fail(SBAMD_ERR_INVALID, "lp_distance takes one PCF on each side");
      if (f->dtype != g->dtype)
        fail(SBAMD_ERR_INVALID, "f and g must have the same dtype");
      if (!(p >= 1.0))
        fail(SBAMD_ERR_INVALID, "p must be >= 1.");
      const int32_t prec = f->dtype == SBAMD_F64 ? SBAMD_PREC_FP64 : SBAMD_PREC_FP32C;
      sbamd_set fs, gs;
      make_set(ctx, *f, prec, fs);
      make_set(ctx, *g, prec, gs);
      sbamd_plan* pl = nullptr;
      const int32_t op = p == 1.0 ? SBAMD_OP_L1 : SBAMD_OP_LP;
      int rc = sbamd_plan_create(ctx, &fs, &gs, op, p, SBAMD_OUT_DENSE, nullptr, &pl);
      if (rc != SBAMD_OK)
        fail(rc, g_last_error);
      std::unique_ptr<sbamd_plan> plan(pl);
      DeviceBuffer dout(sizeof(double));
      execute_range(*plan, dout.ptr, SBAMD_F64, nullptr, 0, work_units(*plan),
                    ctx->stream);
      SBAMD_HIP(hipStreamSynchronize(ctx->stream));
      double r = 0.0;
      SBAMD_HIP(hipMemcpy(&r, dout.ptr, sizeof(double), hipMemcpyDeviceToHost));
      *out = r;
    });
  }

  int sbamd_lp_norm(sbamd_ctx* ctx, const sbamd_pcf_set* set, double p, int32_t precision,
                    void* out_host, int32_t out_dtype)
  {
    return guarded([&] {
      need(ctx, "ctx");
      need(set, "set");
      check_out_dtype(out_dtype);
      if (!(p >= 1.0))
        fail(SBAMD_ERR_INVALID, "p must be >= 1.");
      validate_set(*set);
      if (set->n == 0)
        return;
      need(out_host, "out_host");
      sbamd_set s;
      make_set(ctx, *set, precision, s);
      const size_t bytes = static_cast<size_t>(s.n) * (out_dtype == SBAMD_F64 ? 8 : 4);
      DeviceBuffer dout(bytes);
      int rc = sbamd_set_lp_norm(&s, p, dout.ptr, out_dtype, ctx->stream);
      if (rc != SBAMD_OK)
        fail(rc, g_last_error);
      SBAMD_HIP(hipStreamSynchronize(ctx->stream));
      SBAMD_HIP(hipMemcpy(out_host, dout.ptr, bytes, hipMemcpyDeviceToHost));
    });
  }

  int sbamd_iterate_rectangles(sbamd_ctx* ctx, const sbamd_pcf_set* f, const sbamd_pcf_set* g,
                               double a, double b, void* out, int64_t cap, int64_t* count)
  {
    return guarded([&] {
      need(ctx, "ctx");
      need(f, "f");
      need(g, "g");
      need(count, "count");
      if (f->n != 1 || g->n != 1)
        fail(SBAMD_ERR_INVALID, "iterate_rectangles takes one PCF on each side");
      if (f->dtype != g->dtype)
        fail(SBAMD_ERR_INVALID, "f and g must have the same dtype");
      validate_set(*f);
      validate_set(*g);
      if (cap > 0)
        need(out, "out");
      use_device(ctx);
      const size_t eb = f->dtype == SBAMD_F64 ? 8 : 4;
      // Empty PCFs are the zero PCF [(0, 0)].
      const double zero64 = 0.0;
      const float zero32 = 0.f;
      auto pts = [&](const sbamd_pcf_set* s, const void*& t, const void*& v) {
        int64_t m = s->offsets[1] - s->offsets[0];
        if (m == 0)
        {
          t = f->dtype == SBAMD_F64 ? static_cast<const void*>(&zero64) : static_cast<const void*>(&zero32);
          v = t;
          return int64_t(1);
        }
        t = static_cast<const char*>(s->t) + s->offsets[0] * eb;
        v = static_cast<const char*>(s->v) + s->offsets[0] * eb;
        return m;
      };
      const void *ft, *fv, *gt, *gv;
      const int64_t mf = pts(f, ft, fv);
      const int64_t mg = pts(g, gt, gv);
      DeviceBuffer dft(mf * eb), dfv(mf * eb), dgt(mg * eb), dgv(mg * eb);
      DeviceBuffer dout(static_cast<size_t>(std::max<int64_t>(cap, 1)) * 4 * eb), dcnt(sizeof(int64_t));
      SBAMD_HIP(hipMemcpy(dft.ptr, ft, mf * eb, hipMemcpyHostToDevice));
      SBAMD_HIP(hipMemcpy(dfv.ptr, fv, mf * eb, hipMemcpyHostToDevice));
      SBAMD_HIP(hipMemcpy(dgt.ptr, gt, mg * eb, hipMemcpyHostToDevice));
      SBAMD_HIP(hipMemcpy(dgv.ptr, gv, mg * eb, hipMemcpyHostToDevice));
      if (f->dtype == SBAMD_F64)
        launch_rectangles(dft.as<double>(), dfv.as<double>(), mf, dgt.as<double>(), dgv.as<double>(),
                          mg, a, b, dout.as<double>(), cap, dcnt.as<int64_t>(), ctx->stream);
      else
        launch_rectangles_f32(dft.as<float>(), dfv.as<float>(), mf, dgt.as<float>(), dgv.as<float>(),
                              mg, static_cast<float>(a), static_cast<float>(b), dout.as<float>(), cap,
                              dcnt.as<int64_t>(), ctx->stream);
      SBAMD_HIP(hipStreamSynchronize(ctx->stream));
      int64_t n = 0;
      SBAMD_HIP(hipMemcpy(&n, dcnt.ptr, sizeof(int64_t), hipMemcpyDeviceToHost));
      *count = n;
      const int64_t w = std::min(n, cap);
      if (w > 0)
        SBAMD_HIP(hipMemcpy(out, dout.ptr, static_cast<size_t>(w) * 4 * eb, hipMemcpyDeviceToHost));
    });
  }

  int sbamd_count_segments(const sbamd_pcf_set* rows, const sbamd_pcf_set* cols,
                           int32_t include_diagonal, uint64_t* segments)
  {
    return guarded([&] {
      need(rows, "rows");
      need(segments, "segments");
      validate_set(*rows);
      auto mr = record_counts(*rows);
      uint64_t M = 0;
      for (int64_t x : mr)
        M += static_cast<uint64_t>(x);
      const uint64_t n = static_cast<uint64_t>(rows->n);
      if (cols)
      {
        validate_set(*cols);
        auto mc = record_counts(*cols);
        uint64_t M2 = 0;
        for (int64_t x : mc)
          M2 += static_cast<uint64_t>(x);
        const uint64_t n2 = static_cast<uint64_t>(cols->n);
        *segments = n2 * M + n * M2 - n * n2;
      }
      else if (include_diagonal)
        *segments = n > 0 ? n * M - (n > 1 ? tri(n) : 0) : 0;
      else
        *segments = n > 1 ? (n - 1) * M - tri(n) : 0;
    });
  }

  int sbamd_queue_order_host(int64_t n_tiles, const int32_t* tiles, const int64_t* work, int32_t tile_rows,
                             int32_t tile_cols, int32_t n_queues, int64_t* order, int64_t* queue_off)
  {
    return guarded([&] {
      need(order, "order");
      need(queue_off, "queue_off");
      if (n_tiles < 0 || (n_tiles > 0 && (!tiles || !work)))
        fail(SBAMD_ERR_INVALID, "bad tile list");
      if (n_queues < 1 || n_queues > 64 || tile_rows < 1 || tile_cols < 1)
        fail(SBAMD_ERR_INVALID, "bad queue count or tile shape");
      // the tiles carry their input index in c0's place through the reorder
      std::vector<Tile> t(static_cast<size_t>(n_tiles));
      std::vector<int64_t> w(work, work + n_tiles);
      for (int64_t k = 0; k < n_tiles; ++k)
      {
        t[k].r0 = tiles[3 * k + 0];
        t[k].c0 = tiles[3 * k + 1];
        t[k].prob = static_cast<uint16_t>(tiles[3 * k + 2]);
        t[k].nr = static_cast<int16_t>(tile_rows);
        t[k].nc = static_cast<int16_t>(tile_cols);
        t[k].flags = 0;
      }
      std::vector<Tile> moved(t);
      std::vector<int64_t> qoff(static_cast<size_t>(n_queues) + 1);
      xcd_queue_order(moved, 0, moved.size(), tile_rows, tile_cols, n_queues, w, qoff.data());
      // recover each moved tile's input index: match (prob, r0, c0), unique per tile
      std::map<std::tuple<int, int, int>, int64_t> where;
      for (int64_t k = 0; k < n_tiles; ++k)
        where[std::make_tuple(static_cast<int>(t[k].prob), t[k].r0, t[k].c0)] = k;
      for (int64_t k = 0; k < n_tiles; ++k)
        order[k] = where.at(std::make_tuple(static_cast<int>(moved[k].prob), moved[k].r0, moved[k].c0));
      for (int x = 0; x <= n_queues; ++x)
        queue_off[x] = qoff[x];
    });
  }

  int sbamd_plan_tiles_host(int64_t n_rows, const int64_t* m_rows, int64_t n_cols,
                            const int64_t* m_cols, int32_t tile_rows, int32_t tile_cols,
                            int32_t* tiles, int64_t cap, int64_t* n_tiles, int64_t* perm_rows)
  {
    return guarded([&] {
      need(n_tiles, "n_tiles");
      if (n_rows < 0 || (n_rows > 0 && !m_rows))
        fail(SBAMD_ERR_INVALID, "bad row counts");
      std::vector<int64_t> mr(m_rows, m_rows + n_rows);
      std::vector<int64_t> pr;
      sort_by_records(mr, pr);
      std::vector<int64_t> mrs(mr.size());
      for (size_t q = 0; q < pr.size(); ++q)
        mrs[q] = mr[pr[q]];
      std::vector<Tile> out;
      if (m_cols)
      {
        std::vector<int64_t> mc(m_cols, m_cols + n_cols);
        std::vector<int64_t> pc;
        sort_by_records(mc, pc);
        std::vector<int64_t> mcs(mc.size());
        for (size_t q = 0; q < pc.size(); ++q)
          mcs[q] = mc[pc[q]];
        check_shape(tile_rows, tile_cols, 64, 2);
        build_tiles(mrs, &mcs, tile_rows, tile_cols, false, 0, std::numeric_limits<int64_t>::max(), out);
      }
      else
      {
        check_shape(tile_rows, tile_cols, 64, 2);
        build_tiles(mrs, nullptr, tile_rows, tile_cols, false, 0, std::numeric_limits<int64_t>::max(), out);
      }
      *n_tiles = static_cast<int64_t>(out.size());
      if (tiles)
        for (int64_t k = 0; k < std::min<int64_t>(cap, static_cast<int64_t>(out.size())); ++k)
        {
          tiles[4 * k + 0] = out[k].r0;
          tiles[4 * k + 1] = out[k].c0;
          tiles[4 * k + 2] = out[k].nr;
          tiles[4 * k + 3] = out[k].nc | ((out[k].flags & TILE_DIAG) ? (1 << 16) : 0);
        }
      if (perm_rows)
        for (size_t q = 0; q < pr.size(); ++q)
          perm_rows[q] = pr[q];
    });
  }
}
