// dgv_aio.cpp -- A-array files in the reference's layout, host side.
//
//   WriteAarraysDGV                 readwrite.f90:1459-1484   one file, two records
//   ReadAarraysDGV                  readwrite.f90:1066-1118
//   ReadAarraysChnksDGV             readwrite.f90:1130-1243   a set of chunk files, all records
//   ReadAarraysChnksOfstNrecDGV     readwrite.f90:1261-1449   numrec records from chunk fstchnk on, after ofst
//   ScanAarrsChnks4Acapphi          readwrite.f90:2292-2365   records per chunk
//   MakeBaseNameAoperChnksDGV       readwrite.f90:2224-2281   chunk file names
//
// A file is Fortran unformatted sequential: record 1 = (mm, nn) as two int32,
// record 2 = A_capphi(mm) int32, A(nn) float64, A_xi(nn), A_xi1(nn), A_phi(nn)
// int32.  Each record is framed by its length in bytes as a 4-byte integer
// before and after the payload.  Record 2 exceeds 2^31 - 1 bytes from the 21^3
// grid on, and compilers frame such records differently:
//   * Intel Fortran and gfortran split the record into subrecords of at most
//     2^31 - 9 bytes, each framed by its own length; the leading marker of every
//     subrecord but the last is negated (the record continues), and so is the
//     trailing marker of every subrecord but the first;
//   * LLVM Flang (amdflang) writes one record whose markers hold the length
//     modulo 2^32 (read as an unsigned 32-bit integer).
// The reader recognises either (the payload length follows from mm and nn);
// the writer takes the framing as an argument.
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "dgv_internal.h"

using namespace dgv;

namespace {

// largest subrecord payload of the Intel / gfortran convention
int64_t g_subrecord_max = 2147483639;

struct Part {
  const void* p;
  int64_t n;
};

bool write_marker(std::ofstream& o, int32_t m) {
  o.write(reinterpret_cast<const char*>(&m), 4);
  return (bool)o;
}

// one record made of the given parts, in the given framing
bool write_record(std::ofstream& o, const std::vector<Part>& parts, int framing) {
  int64_t len = 0;
  for (const Part& p : parts) len += p.n;
  if (framing == DGV_FRAMING_UNSIGNED32 || len <= g_subrecord_max) {
    const int32_t m = (int32_t)(uint32_t)(uint64_t)len;   // modulo 2^32 for the unsigned framing
    if (!write_marker(o, m)) return false;
    for (const Part& p : parts) o.write(static_cast<const char*>(p.p), (std::streamsize)p.n);
    return write_marker(o, m);
  }
  // subrecords: walk the parts in slices of at most g_subrecord_max bytes
  size_t pi = 0;
  int64_t off = 0, left = len;
  bool first = true;
  while (left > 0) {
    const int64_t sub = left < g_subrecord_max ? left : g_subrecord_max;
    const bool last = sub == left;
    if (!write_marker(o, last ? (int32_t)sub : -(int32_t)sub)) return false;
    int64_t need = sub;
    while (need > 0) {
      const int64_t take = std::min<int64_t>(need, parts[pi].n - off);
      o.write(static_cast<const char*>(parts[pi].p) + off, (std::streamsize)take);
      need -= take;
      off += take;
      if (off == parts[pi].n) { ++pi; off = 0; }
    }
    if (!write_marker(o, first ? (int32_t)sub : -(int32_t)sub)) return false;
    first = false;
    left -= sub;
  }
  return (bool)o;
}

// Sequential reader of one record's payload whose total length `len` is
// known (from the header record): it recognises the framing at the leading
// marker and then reads across subrecord boundaries transparently.
class RecordReader {
 public:
  RecordReader(std::ifstream& in, int64_t len) : in_(in), len_(len) {}

  // returns an empty string or why the framing was not recognised
  std::string open() {
    if (!rd(&lead_, 4)) return "missing record marker";
    const std::streampos data0 = in_.tellg();
    if (lead_ >= 0 && (int64_t)lead_ == len_) {   // plain record
      mode_ = kPlain;
      left_ = len_;
      return "";
    }
    if ((uint32_t)lead_ == (uint32_t)(uint64_t)len_) {   // LLVM Flang: check the trailer
      in_.seekg(data0 + (std::streamoff)len_);
      int32_t t = 0;
      const bool ok = rd(&t, 4) && t == lead_;
      in_.clear();
      in_.seekg(data0);
      if (ok) {
        mode_ = kPlain;
        left_ = len_;
        return "";
      }
    }
    if (lead_ < 0) {   // Intel / gfortran subrecords
      mode_ = kSub;
      left_ = -(int64_t)lead_;
      sub_ = left_;
      if (left_ > len_) return "subrecord longer than the payload";
      return "";
    }
    return "record length " + std::to_string(lead_) + " does not match the expected " + std::to_string(len_) +
           " bytes";
  }

  // the next n bytes of the payload
  std::string read(void* dst, int64_t n) {
    char* d = static_cast<char*>(dst);
    while (n > 0) {
      if (left_ == 0) {
        if (std::string e = next_subrecord(); !e.empty()) return e;
        continue;
      }
      const int64_t take = std::min(n, left_);
      if (!rd(d, take)) return "truncated record";
      d += take;
      n -= take;
      left_ -= take;
      got_ += take;
    }
    return "";
  }

  // the rest of the payload is not needed: check the framing up to the record's end
  std::string finish() {
    if (got_ + left_ < len_ && mode_ == kPlain) return "record shorter than its payload";
    std::vector<char> skip(1 << 20);
    while (got_ < len_) {
      const int64_t n = std::min<int64_t>(len_ - got_, (int64_t)skip.size());
      if (std::string e = read(skip.data(), n); !e.empty()) return e;
    }
    int32_t t;
    if (!rd(&t, 4)) return "missing trailing record marker";
    if (mode_ == kPlain ? t != lead_ : t != (first_ ? (int32_t)sub_ : -(int32_t)sub_))
      return "bad trailing record marker";
    if (mode_ == kSub && lead_ < 0) return "subrecords shorter than the payload";
    return "";
  }

 private:
  enum { kPlain, kSub };
  std::string next_subrecord() {
    if (mode_ != kSub || lead_ > 0) return "payload longer than the record";
    int32_t t;
    if (!rd(&t, 4) || t != (first_ ? (int32_t)sub_ : -(int32_t)sub_)) return "bad subrecord trailing marker";
    first_ = false;
    if (!rd(&lead_, 4)) return "missing subrecord marker";
    sub_ = lead_ < 0 ? -(int64_t)lead_ : (int64_t)lead_;
    if (got_ + sub_ > len_) return "subrecords longer than the payload";
    left_ = sub_;
    return "";
  }
  bool rd(void* dst, int64_t n) {
    in_.read(static_cast<char*>(dst), (std::streamsize)n);
    return (bool)in_;
  }
  std::ifstream& in_;
  int64_t len_;
  int mode_ = kPlain;
  int32_t lead_ = 0;
  int64_t left_ = 0, got_ = 0, sub_ = 0;
  bool first_ = true;
};

std::string read_record(std::ifstream& in, const std::vector<Part>& parts) {
  int64_t len = 0;
  for (const Part& p : parts) len += p.n;
  RecordReader r(in, len);
  if (std::string e = r.open(); !e.empty()) return e;
  for (const Part& p : parts)
    if (std::string e = r.read(const_cast<void*>(p.p), p.n); !e.empty()) return e;
  return r.finish();
}

std::string trimmed(const char* s) {
  std::string t = s ? s : "";
  const size_t a = t.find_first_not_of(' ');
  if (a == std::string::npos) return "";
  const size_t b = t.find_last_not_of(' ');
  return t.substr(a, b - a + 1);
}

std::string chunk_path(const dgv_grid_desc& g, const char* dir, const char* base, int32_t chunk, int32_t k_c,
                       int32_t N_x) {
  char buf[2048];
  std::snprintf(buf, sizeof(buf), "%s%sch%03d_%dkc%dsu%dsv%dsw%dNXU%dMuUU%dMvVU%dMwWU_Aarrs.dat",
                trimmed(dir).c_str(), trimmed(base).c_str(), chunk, k_c, g.su, g.sv, g.sw, N_x, g.Mu, g.Mv, g.Mw);
  return buf;
}

int read_header(const std::string& path, std::ifstream& in, int64_t& mm, int64_t& nn, std::string& err) {
  in.open(path, std::ios::binary);
  if (!in) {
    err = "cannot open " + path;
    return DGV_ERR_IO;
  }
  int32_t h[2];
  const std::string r = read_record(in, {{h, 8}});
  if (!r.empty() || h[0] < 0 || h[1] < 0) {
    err = path + ": bad header record (" + (r.empty() ? std::string("negative sizes") : r) + ")";
    return DGV_ERR_IO;
  }
  mm = h[0];
  nn = h[1];
  return DGV_OK;
}

struct Arrays {
  std::vector<int32_t> capphi, xi, xi1, phi;
  std::vector<double> A;
};

int read_file(const std::string& path, Arrays& a, std::string& err) {
  std::ifstream in;
  int64_t mm, nn;
  if (int rc = read_header(path, in, mm, nn, err)) return rc;
  try {
    a.capphi.resize(mm);
    a.A.resize(nn);
    a.xi.resize(nn);
    a.xi1.resize(nn);
    a.phi.resize(nn);
  } catch (const std::bad_alloc&) {
    err = "out of host memory for " + path;
    return DGV_ERR_NOMEM;
  }
  const std::string r = read_record(in, {{a.capphi.data(), 4 * mm}, {a.A.data(), 8 * nn}, {a.xi.data(), 4 * nn},
                                         {a.xi1.data(), 4 * nn}, {a.phi.data(), 4 * nn}});
  if (!r.empty()) {
    err = path + ": " + r;
    return DGV_ERR_IO;
  }
  return DGV_OK;
}

int write_file(const std::string& path, const int32_t* capphi, int64_t mm, const double* A, const int32_t* xi,
               const int32_t* xi1, const int32_t* phi, int64_t nn, int framing, std::string& err) {
  if (mm > INT32_MAX || nn > INT32_MAX) {
    err = "more than 2^31 - 1 nodes or records: the header record holds int32 sizes";
    return DGV_ERR_UNSUPPORTED;
  }
  std::ofstream o(path, std::ios::binary | std::ios::trunc);
  if (!o) {
    err = "cannot open " + path + " for writing";
    return DGV_ERR_IO;
  }
  const int32_t h[2] = {(int32_t)mm, (int32_t)nn};
  if (!write_record(o, {{h, 8}}, framing) ||
      !write_record(o, {{capphi, 4 * mm}, {A, 8 * nn}, {xi, 4 * nn}, {xi1, 4 * nn}, {phi, 4 * nn}}, framing)) {
    err = "write failed: " + path;
    return DGV_ERR_IO;
  }
  return DGV_OK;
}

int global_fail(int code, const std::string& m) {
  set_global_error(m);
  return code;
}

}  // namespace

namespace dgv {
int get_all_records(const dgv_ctx* ctx, std::vector<int32_t>& capphi, std::vector<double>& A,
                    std::vector<int32_t>& xi, std::vector<int32_t>& xi1, int index_base);
}

extern "C" {

int dgv_set_fortran_subrecord_limit(int64_t bytes) {
  if (bytes < 8 || bytes > 2147483639) return global_fail(DGV_ERR_ARG, "subrecord limit must be in 8 .. 2^31 - 9");
  g_subrecord_max = bytes;
  return DGV_OK;
}

int dgv_aarray_file_header(const char* path, int64_t* mm, int64_t* nn) {
  if (!path || !mm || !nn) return global_fail(DGV_ERR_ARG, "null argument");
  std::ifstream in;
  std::string err;
  if (int rc = read_header(path, in, *mm, *nn, err)) return global_fail(rc, err);
  return DGV_OK;
}

int dgv_aarray_file_read(const char* path, int32_t* capphi, int64_t mm, double* A, int32_t* xi, int32_t* xi1,
                         int32_t* phi, int64_t nn) {
  if (!path || (mm > 0 && !capphi) || (nn > 0 && (!A || !xi || !xi1 || !phi)))
    return global_fail(DGV_ERR_ARG, "null argument");
  std::ifstream in;
  std::string err;
  int64_t fm, fn;
  if (int rc = read_header(path, in, fm, fn, err)) return global_fail(rc, err);
  if (fm != mm || fn != nn) return global_fail(DGV_ERR_ARG, std::string(path) + ": sizes differ from the header");
  const std::string r = read_record(in, {{capphi, 4 * mm}, {A, 8 * nn}, {xi, 4 * nn}, {xi1, 4 * nn}, {phi, 4 * nn}});
  if (!r.empty()) return global_fail(DGV_ERR_IO, std::string(path) + ": " + r);
  return DGV_OK;
}

int dgv_aarray_file_write(const char* path, const int32_t* capphi, int64_t mm, const double* A, const int32_t* xi,
                          const int32_t* xi1, const int32_t* phi, int64_t nn, int32_t framing) {
  if (!path || mm < 0 || nn < 0 || (mm > 0 && !capphi) || (nn > 0 && (!A || !xi || !xi1 || !phi)))
    return global_fail(DGV_ERR_ARG, "null or negative argument");
  if (framing != DGV_FRAMING_SUBRECORDS && framing != DGV_FRAMING_UNSIGNED32)
    return global_fail(DGV_ERR_ARG, "framing must be DGV_FRAMING_SUBRECORDS or DGV_FRAMING_UNSIGNED32");
  std::string err;
  if (int rc = write_file(path, capphi, mm, A, xi, xi1, phi, nn, framing, err)) return global_fail(rc, err);
  return DGV_OK;
}

// ScanAarrsChnks4Acapphi: chunks_Act[i] = sum of chunk i's A_capphi
int dgv_aarray_chunks_scan(const dgv_grid_desc* grid, const char* dir, const char* base, int32_t k_c, int32_t N_x,
                           int32_t nchnks, int64_t* chunks_Act) {
  if (!grid || nchnks < 1 || !chunks_Act) return global_fail(DGV_ERR_ARG, "bad argument");
  for (int32_t i = 0; i < nchnks; ++i) {
    std::ifstream in;
    int64_t mm, nn;
    std::string err;
    const std::string path = chunk_path(*grid, dir, base, i, k_c, N_x);
    if (int rc = read_header(path, in, mm, nn, err)) return global_fail(rc, err);
    std::vector<int32_t> cap(mm);
    // A_capphi is the first part of record 2: only it is read
    RecordReader r(in, 4 * mm + 20 * nn);
    std::string e = r.open();
    if (e.empty()) e = r.read(cap.data(), 4 * mm);
    if (!e.empty()) return global_fail(DGV_ERR_IO, path + ": " + e);
    int64_t s = 0;
    for (int32_t c : cap) s += c;
    chunks_Act[i] = s;
  }
  return DGV_OK;
}

// ReadAarraysChnksOfstNrecDGV (readwrite.f90:1261-1449): numrec records
// starting after the first ofst records of chunk fstchnk (1-based), continuing
// into chunks fstchnk + 1 .. nchnks; capphi[mm] receives the per-node counts of
// the records read, in the reference's bookkeeping.  *nread = records read.
// ReadAarraysChnksDGV is fstchnk = 1, ofst = 0, numrec = all records.
int dgv_aarray_chunks_read(const dgv_grid_desc* grid, const char* dir, const char* base, int32_t k_c, int32_t N_x,
                           int32_t fstchnk, int64_t ofst, int64_t numrec, int32_t nchnks, int32_t* capphi,
                           int64_t mm_cap, double* A, int32_t* xi, int32_t* xi1, int32_t* phi, int64_t nrec_cap,
                           int64_t* nread) {
  if (!grid || fstchnk < 1 || fstchnk > nchnks || ofst < 0 || numrec < 0 || !capphi || !nread ||
      (numrec > 0 && (!A || !xi || !xi1 || !phi)) || nrec_cap < numrec)
    return global_fail(DGV_ERR_ARG, "bad argument");
  std::string err;
  Arrays c;
  std::string path = chunk_path(*grid, dir, base, fstchnk - 1, k_c, N_x);
  if (int rc = read_file(path, c, err)) return global_fail(rc, err);
  const int64_t mm = (int64_t)c.capphi.size(), nn = (int64_t)c.A.size();
  if (mm_cap < mm) return global_fail(DGV_ERR_ARG, "capphi capacity smaller than the chunk's node count");
  std::fill(capphi, capphi + mm, 0);
  // the first chunk: skip ofst records, take up to numrec
  int64_t need = std::min<int64_t>(nn - ofst, numrec);
  if (need < 0) return global_fail(DGV_ERR_INDEX, "offset past the first chunk's records");
  std::copy(c.A.begin() + ofst, c.A.begin() + ofst + need, A);
  std::copy(c.xi.begin() + ofst, c.xi.begin() + ofst + need, xi);
  std::copy(c.xi1.begin() + ofst, c.xi1.begin() + ofst + need, xi1);
  std::copy(c.phi.begin() + ofst, c.phi.begin() + ofst + need, phi);
  int64_t act = need;
  // A_capphi of what was taken: the node in which the offset falls first
  int64_t i = 1, cum = c.capphi[0];
  while (ofst > cum && i <= mm) {   // the first node whose records reach past the offset
    ++i;
    if (i <= mm) cum += c.capphi[i - 1];
  }
  if (i > mm) return global_fail(DGV_ERR_INDEX, "no records in the first chunk after the offset");
  if (cum - ofst < need) {
    capphi[i - 1] = (int32_t)(cum - ofst);
    need -= cum - ofst;
  } else {
    capphi[i - 1] = (int32_t)need;
    need = 0;
  }
  while (need > 0 && i < mm) {
    ++i;
    if (c.capphi[i - 1] >= need) {
      capphi[i - 1] = (int32_t)need;
      need = 0;
    } else {
      capphi[i - 1] = c.capphi[i - 1];
      need -= c.capphi[i - 1];
    }
  }
  if (i >= mm && need > 0) return global_fail(DGV_ERR_INDEX, "records of the first chunk do not add up");
  // the following chunks, no offset
  for (int32_t j = fstchnk + 1; act < numrec && j <= nchnks; ++j) {
    path = chunk_path(*grid, dir, base, j - 1, k_c, N_x);
    if (int rc = read_file(path, c, err)) return global_fail(rc, err);
    if ((int64_t)c.capphi.size() != mm) return global_fail(DGV_ERR_IO, path + ": node count differs between chunks");
    const int64_t n2 = (int64_t)c.A.size();
    need = std::min<int64_t>(n2, numrec - act);
    std::copy(c.A.begin(), c.A.begin() + need, A + act);
    std::copy(c.xi.begin(), c.xi.begin() + need, xi + act);
    std::copy(c.xi1.begin(), c.xi1.begin() + need, xi1 + act);
    std::copy(c.phi.begin(), c.phi.begin() + need, phi + act);
    act += need;
    int64_t k = 0;
    while (need > 0 && k < mm) {
      ++k;
      if (c.capphi[k - 1] >= need) {
        capphi[k - 1] += (int32_t)need;
        need = 0;
      } else {
        capphi[k - 1] += c.capphi[k - 1];
        need -= c.capphi[k - 1];
      }
    }
    if (k >= mm && need > 0) return global_fail(DGV_ERR_INDEX, "records of a chunk do not add up");
  }
  *nread = act;
  return DGV_OK;
}

// The context's A-array as a chunk set: records in the reference order, at most
// max_records per chunk, each chunk's A_capphi counting its own records.
int dgv_ctx_write_aarray_chunks(const dgv_ctx* ctx, const char* dir, const char* base, int32_t k_c, int32_t N_x,
                                int64_t max_records, int32_t framing, int32_t* nchunks) {
  if (!ctx) return global_fail(DGV_ERR_ARG, "null context");
  dgv_ctx* c = const_cast<dgv_ctx*>(ctx);
  if (!ctx->has_a) return set_error(c, DGV_ERR_STATE, "no A-array set");
  if (max_records < 1 || !nchunks) return set_error(c, DGV_ERR_ARG, "max_records must be >= 1");
  if (framing != DGV_FRAMING_SUBRECORDS && framing != DGV_FRAMING_UNSIGNED32)
    return set_error(c, DGV_ERR_ARG, "unknown framing");
  std::vector<int32_t> capphi, xi, xi1;
  std::vector<double> A;
  if (get_all_records(ctx, capphi, A, xi, xi1, 1)) return set_error(c, DGV_ERR_HIP, "A-array export failed");
  const int64_t mm = (int64_t)capphi.size(), nn = (int64_t)A.size();
  std::vector<int32_t> phi(nn);
  for (int64_t k = 0, r = 0; k < mm; ++k)
    for (int32_t j = 0; j < capphi[k]; ++j, ++r) phi[r] = (int32_t)(k + 1);
  const int64_t nch = std::max<int64_t>(1, (nn + max_records - 1) / max_records);
  if (nch > 1000) return set_error(c, DGV_ERR_ARG, "more than 1000 chunks (names carry three digits)");
  dgv_grid_desc gd{ctx->g.Mu, ctx->g.Mv, ctx->g.Mw, ctx->g.su, ctx->g.sv, ctx->g.sw,
                   ctx->g.box[0], ctx->g.box[1], ctx->g.box[2], ctx->g.box[3], ctx->g.box[4], ctx->g.box[5]};
  for (int64_t ch = 0; ch < nch; ++ch) {
    const int64_t r0 = ch * max_records, r1 = std::min(nn, r0 + max_records);
    std::vector<int32_t> cap(mm, 0);
    for (int64_t r = r0; r < r1; ++r) ++cap[phi[r] - 1];
    std::string err;
    const std::string path = chunk_path(gd, dir, base, (int32_t)ch, k_c, N_x);
    if (int rc = write_file(path, cap.data(), mm, A.data() + r0, xi.data() + r0, xi1.data() + r0, phi.data() + r0,
                            r1 - r0, framing, err))
      return set_error(c, rc, err);
  }
  *nchunks = (int32_t)nch;
  return DGV_OK;
}

// ReadAarraysChnksDGV into the context: every record of chunks 0 .. nchnks-1
int dgv_ctx_read_aarray_chunks(dgv_ctx* ctx, const char* dir, const char* base, int32_t k_c, int32_t N_x,
                               int32_t nchnks) {
  if (!ctx) return global_fail(DGV_ERR_ARG, "null context");
  if (nchnks < 1) return set_error(ctx, DGV_ERR_ARG, "nchnks must be >= 1");
  dgv_grid_desc gd{ctx->g.Mu, ctx->g.Mv, ctx->g.Mw, ctx->g.su, ctx->g.sv, ctx->g.sw,
                   ctx->g.box[0], ctx->g.box[1], ctx->g.box[2], ctx->g.box[3], ctx->g.box[4], ctx->g.box[5]};
  std::vector<int64_t> act(nchnks);
  if (int rc = dgv_aarray_chunks_scan(&gd, dir, base, k_c, N_x, nchnks, act.data()))
    return set_error(ctx, rc, dgv_last_error(nullptr));
  int64_t total = 0;
  for (int64_t a : act) total += a;
  std::vector<int32_t> capphi(ctx->g.N), xi(total), xi1(total), phi(total);
  std::vector<double> A(total);
  int64_t nread = 0;
  if (int rc = dgv_aarray_chunks_read(&gd, dir, base, k_c, N_x, 1, 0, total, nchnks, capphi.data(), ctx->g.N,
                                      A.data(), xi.data(), xi1.data(), phi.data(), total, &nread))
    return set_error(ctx, rc, dgv_last_error(nullptr));
  if (nread != total) return set_error(ctx, DGV_ERR_IO, "chunk set shorter than its A_capphi");
  return build_pairblock_from_records(ctx, capphi.data(), ctx->g.N, A.data(), xi.data(), xi1.data(), total, 1);
}

}  // extern "C"
