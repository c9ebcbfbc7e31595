// The reference's binary file format (include/sbear/io.hpp:27-158,
// io/io_stream_base.hpp, io/tensor_io.hpp, io/pcf_io.hpp, io/point_io.hpp,
// io/compressed_matrix_io.hpp): parsing and indexing of whole file images in
// memory, and header and PCF-element writers. The C-ABI wrappers and the
// device set built straight from a file image are in api.cpp.
//
// Layout (all little-endian, no padding):
//   header  : "\1MPCF", 'e', int32 format version (1 or 2),
//             u64-length-prefixed strings: project version, build date and,
//             from version 2, platform
//   u32 kind: 1 = one tensor, 2 = one object
//   int32 base format, int32 sub format (element type, width in bits)
//   tensor  : u64 rank, then (u64 extent, u64 stride) per dimension, strides
//             row-major in elements; then the elements in row-major order
//   object  : one element
//   elements: numbers as raw bytes (bool: 1 byte); a PCF as u64 point count
//             and (time, value) per point; a compressed matrix as u64 n and
//             its storage_count() values.
#include "sbamd_internal.hpp"

#include <cstring>
#include <string>

namespace sbamd
{
  namespace
  {
    const char HEADER_ID[5] = {'\1', 'M', 'P', 'C', 'F'};
    constexpr int32_t FORMAT_VERSION = 2;
    constexpr uint64_t MAX_STRING = 32767;
    constexpr int IO_MAX_RANK = 32;

    const char* const SHORT_READ = "Incorrect number of bytes returned (file may be corrupted)";

    struct Reader
    {
      const unsigned char* p;
      uint64_t len;
      uint64_t pos;

      uint64_t left() const { return len - pos; }

      template <typename T>
      T get()
      {
        if (left() < sizeof(T))
          fail(SBAMD_ERR_RUNTIME, SHORT_READ);
        T v;
        std::memcpy(&v, p + pos, sizeof(T));
        pos += sizeof(T);
        return v;
      }

      std::string raw(uint64_t n)
      {
        if (left() < n)
          fail(SBAMD_ERR_RUNTIME, "Unexpected EOF");
        std::string s(reinterpret_cast<const char*>(p + pos), n);
        pos += n;
        return s;
      }

      std::string str()
      {
        const uint64_t n = get<uint64_t>();
        if (n > MAX_STRING)
          fail(SBAMD_ERR_RUNTIME, "Read string of length " + std::to_string(n) + " which exceeds the " +
                                      std::to_string(MAX_STRING) +
                                      " limit (this is likely a corrupted file).");
        return raw(n);
      }
    };

    // Element type codes (tensor_io.hpp:92-121).
    bool known_format(int32_t base, int32_t sub)
    {
      switch (base)
      {
      case 1:
      case 2:
      case 3:
      case 100:
      case 101:
      case 1000:
      case 1100:
      case 1120:
      case 10000:
        return sub == 32 || sub == 64;
      case 4:
        return sub == 8;
      default:
        return false;
      }
    }

    bool object_format(int32_t base) { return base == 100 || base == 101 || base == 1100 || base == 1120 || base == 10000; }

    std::string format_name(uint32_t kind)
    {
      if (kind == SBAMD_IO_TENSOR)
        return "SingleTensor";
      if (kind == SBAMD_IO_OBJECT)
        return "SingleObject";
      return "Unknown(" + std::to_string(kind) + ")";
    }

    std::string dims(const uint64_t* v, int n)
    {
      std::string s = "(";
      for (int k = 0; k < n; ++k)
        s += (k ? ", " : "") + std::to_string(v[k]);
      return s + ")";
    }

    // Row-major strides (in elements) of a shape, as the reference's Tensor
    // constructor lays them out.
    void row_major(const uint64_t* shape, int ndim, uint64_t* strides)
    {
      uint64_t s = 1;
      for (int k = ndim - 1; k >= 0; --k)
      {
        strides[k] = s;
        s *= shape[k];
      }
    }

    void parse(const void* buf, uint64_t len, sbamd_io_info& info)
    {
      std::memset(&info, 0, sizeof(info));
      Reader r{static_cast<const unsigned char*>(buf), len, 0};
      if (r.raw(5) != std::string(HEADER_ID, 5))
        fail(SBAMD_ERR_RUNTIME, "Unrecognized file format.");
      if (r.raw(1) != "e")
        fail(SBAMD_ERR_RUNTIME, "Data were saved on a platform with different endianness and cannot be "
                                "read on this platform. If you need to move data in this way, please "
                                "file an issue.");
      const int32_t version = r.get<int32_t>();
      if (version < 1 || version > FORMAT_VERSION)
        fail(SBAMD_ERR_RUNTIME, "Input file has format version " + std::to_string(version) +
                                    ". This version of stablebear reads format versions 1 through " +
                                    std::to_string(FORMAT_VERSION) + ".");
      (void)r.str(); // project version
      (void)r.str(); // build date
      if (version >= 2)
        (void)r.str(); // platform
      info.version = version;
      info.kind = r.get<uint32_t>();
      if (info.kind != SBAMD_IO_TENSOR && info.kind != SBAMD_IO_OBJECT)
        fail(SBAMD_ERR_RUNTIME, "Unknown format type " + format_name(info.kind));
      info.base_format = r.get<int32_t>();
      info.sub_format = r.get<int32_t>();
      const std::string what = info.kind == SBAMD_IO_TENSOR ? "tensor" : "object";
      if (!known_format(info.base_format, info.sub_format) ||
          (info.kind == SBAMD_IO_OBJECT && !object_format(info.base_format)))
        fail(SBAMD_ERR_RUNTIME, "Unhandled " + what + " type (" + std::to_string(info.base_format) + ", " +
                                    std::to_string(info.sub_format) + ")");
      if (info.kind == SBAMD_IO_OBJECT)
      {
        info.ndim = 0;
        info.elements = 1;
      }
      else
      {
        const uint64_t rank = r.get<uint64_t>();
        if (rank > static_cast<uint64_t>(IO_MAX_RANK))
          fail(SBAMD_ERR_RUNTIME, "Tensor of rank " + std::to_string(rank) + " (this build reads ranks up to " +
                                      std::to_string(IO_MAX_RANK) + ")");
        info.ndim = static_cast<int32_t>(rank);
        uint64_t strides[IO_MAX_RANK], expect[IO_MAX_RANK];
        for (int k = 0; k < info.ndim; ++k)
        {
          info.shape[k] = r.get<uint64_t>();
          strides[k] = r.get<uint64_t>();
        }
        row_major(info.shape, info.ndim, expect);
        for (int k = 0; k < info.ndim; ++k)
          if (strides[k] != expect[k])
            fail(SBAMD_ERR_RUNTIME, "Incorrect strides in saved data (expected " + dims(expect, info.ndim) +
                                        " but got " + dims(strides, info.ndim) + ")");
        // the reference's Tensor::size(): 0 for rank 0, else the product
        uint64_t e = info.ndim ? 1 : 0;
        for (int k = 0; k < info.ndim; ++k)
        {
          if (info.shape[k] && e > UINT64_MAX / info.shape[k])
            fail(SBAMD_ERR_RUNTIME, "Tensor shape overflows");
          e *= info.shape[k];
        }
        info.elements = e;
      }
      info.payload = r.pos;
    }

    // Byte extent of each element (PCFs and compressed matrices; numbers are
    // fixed-size and contiguous).
    uint64_t index(const void* buf, uint64_t len, const sbamd_io_info& info, uint64_t* counts, uint64_t* starts)
    {
      Reader r{static_cast<const unsigned char*>(buf), len, info.payload};
      if (info.payload > len)
        fail(SBAMD_ERR_RUNTIME, SHORT_READ);
      const uint64_t es = static_cast<uint64_t>(info.sub_format) / 8;
      const int32_t b = info.base_format;
      if (b == 1 || b == 2 || b == 3 || b == 4)
      {
        if (info.elements && r.left() / info.elements < es)
          fail(SBAMD_ERR_RUNTIME, SHORT_READ);
        for (uint64_t e = 0; e < info.elements; ++e)
        {
          if (counts)
            counts[e] = 1;
          if (starts)
            starts[e] = info.payload + e * es;
        }
        return info.payload + info.elements * es;
      }
      if (b == 1000 || b == 10000)
        fail(SBAMD_ERR_INVALID, "point clouds and barcodes are not supported by this build");
      for (uint64_t e = 0; e < info.elements; ++e)
      {
        const uint64_t n = r.get<uint64_t>();
        uint64_t bytes_each, items;
        if (b == 100 || b == 101)
        {
          bytes_each = 2 * es; // (time, value)
          items = n;
        }
        else
        {
          bytes_each = es;
          const uint64_t lim = 1ull << 32;
          if (n >= lim)
            fail(SBAMD_ERR_RUNTIME, SHORT_READ);
          items = b == 1120 ? (n ? n * (n - 1) / 2 : 0) : n * (n + 1) / 2;
        }
        if (items > r.left() / bytes_each)
          fail(SBAMD_ERR_RUNTIME, SHORT_READ);
        if (counts)
          counts[e] = n;
        if (starts)
          starts[e] = r.pos;
        r.pos += items * bytes_each;
      }
      return r.pos;
    }

    std::string build_date() { return std::string(__DATE__) + " " + __TIME__; }

    struct Writer
    {
      unsigned char* out;
      uint64_t cap;
      uint64_t pos = 0;

      void put(const void* src, uint64_t n)
      {
        if (out)
        {
          if (cap - pos < n || pos > cap)
            fail(SBAMD_ERR_INVALID, "output buffer too small");
          std::memcpy(out + pos, src, n);
        }
        pos += n;
      }
      template <typename T>
      void put(T v) { put(&v, sizeof(T)); }
      void str(const std::string& s)
      {
        put<uint64_t>(s.size());
        put(s.data(), s.size());
      }
    };
  }

  void io_parse(const void* buf, uint64_t len, sbamd_io_info& info)
  {
    if (!buf && len)
      fail(SBAMD_ERR_INVALID, "buf is NULL");
    parse(buf, len, info);
  }

  uint64_t io_index(const void* buf, uint64_t len, const sbamd_io_info& info, uint64_t* counts,
                    uint64_t* starts)
  {
    if (!buf && len)
      fail(SBAMD_ERR_INVALID, "buf is NULL");
    return index(buf, len, info, counts, starts);
  }

  uint64_t io_header(uint32_t kind, int32_t base_format, int32_t sub_format, int32_t ndim,
                     const uint64_t* shape, void* out, uint64_t cap)
  {
    if (kind != SBAMD_IO_TENSOR && kind != SBAMD_IO_OBJECT)
      fail(SBAMD_ERR_INVALID, "kind must be SBAMD_IO_TENSOR or SBAMD_IO_OBJECT");
    if (!known_format(base_format, sub_format) || (kind == SBAMD_IO_OBJECT && !object_format(base_format)))
      fail(SBAMD_ERR_INVALID, "unknown element format (" + std::to_string(base_format) + ", " +
                                  std::to_string(sub_format) + ")");
    if (kind == SBAMD_IO_TENSOR && (ndim < 0 || ndim > IO_MAX_RANK || (ndim > 0 && !shape)))
      fail(SBAMD_ERR_INVALID, "tensor rank must be in [0, 32] with a shape");
    Writer w{static_cast<unsigned char*>(out), cap};
    w.put(HEADER_ID, 5);
    w.put("e", 1);
    w.put<int32_t>(FORMAT_VERSION);
    w.str(std::string("stablebear_amd ") + sbamd_version());
    w.str(build_date());
    w.str("linux-x86_64");
    w.put<uint32_t>(kind);
    w.put<int32_t>(base_format);
    w.put<int32_t>(sub_format);
    if (kind == SBAMD_IO_TENSOR)
    {
      uint64_t strides[IO_MAX_RANK];
      row_major(shape, ndim, strides);
      w.put<uint64_t>(static_cast<uint64_t>(ndim));
      for (int k = 0; k < ndim; ++k)
      {
        w.put<uint64_t>(shape[k]);
        w.put<uint64_t>(strides[k]);
      }
    }
    return w.pos;
  }

  uint64_t io_write_pcfs(int64_t n, const int64_t* offsets, const void* t, const void* v, int32_t dtype,
                         void* out, uint64_t cap)
  {
    check_set_offsets(n, offsets, dtype);
    const uint64_t es = dtype == SBAMD_F64 ? 8 : 4;
    if (n > 0 && offsets[n] > 0 && (!t || !v))
      fail(SBAMD_ERR_INVALID, "PCF set time/value arrays are NULL");
    Writer w{static_cast<unsigned char*>(out), cap};
    const unsigned char* tb = static_cast<const unsigned char*>(t);
    const unsigned char* vb = static_cast<const unsigned char*>(v);
    for (int64_t i = 0; i < n; ++i)
    {
      const int64_t b = offsets[i], e = offsets[i + 1];
      w.put<uint64_t>(static_cast<uint64_t>(e - b));
      for (int64_t k = b; k < e; ++k)
      {
        w.put(tb + k * es, es);
        w.put(vb + k * es, es);
      }
    }
    return w.pos;
  }
}
