// dgv_io.cpp -- the driver's solution checkpoint and file names, host side.
//
//   MakeBaseNameDGV       readwrite.f90:1505-1560   experiment base name (uniform meshes)
//   WriteSol_DGV          readwrite.f90:1756-1796   <base>_time<suff>_sol.dat
//   WriteSol_init_DGV     readwrite.f90:1810-1850   the same file with f alone
//   RestartSolutionDGV    readwrite.f90:2389-2513   reads it back
//
// The file is Fortran unformatted sequential: every record is framed by its
// byte length as a 4-byte integer before and after the payload.  Records:
//   1: e_time, dt            (2 x float64)
//   2: rkmts                 (int32)
//   3: size(f)               (int32)
//   4: f, f1..f_{k-1}, frhs1..frhs_{k-1}   (float64, k = rkmts; f alone for k = 1)
// The history is passed as the steppers hold it: hist = f1..f4, frhs1..frhs5,
// nine arrays of n doubles.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "dgv_internal.h"

using namespace dgv;

namespace {

std::string base_name(const dgv_grid_desc& g, const char* dir, const char* base, int32_t k_c, int32_t N_x) {
  // MakeBaseNameDGV: dir // base // k_c"kc" su"su" sv"sv" sw"sw" N"N" XU Mu"Mu" UU Mv"Mv" VU Mw"Mw" WU
  char buf[1024];
  std::snprintf(buf, sizeof(buf), "%s%s%dkc%dsu%dsv%dsw%dNXU%dMuUU%dMvVU%dMwWU", dir ? dir : "", base ? base : "", k_c,
                g.su, g.sv, g.sw, N_x, g.Mu, g.Mv, g.Mw);
  return buf;
}

// trim(adjustl(s))
std::string trimmed(const char* s) {
  std::string t = s ? s : "";
  const size_t a = t.find_first_not_of(' ');
  if (a == std::string::npos) return "";
  const size_t b = t.find_last_not_of(' ');
  return t.substr(a, b - a + 1);
}

bool put_record(std::ofstream& o, const std::vector<std::pair<const void*, size_t>>& parts) {
  size_t len = 0;
  for (auto& p : parts) len += p.second;
  if (len > (size_t)INT32_MAX) return false;
  const int32_t l = (int32_t)len;
  o.write(reinterpret_cast<const char*>(&l), 4);
  for (auto& p : parts) o.write(static_cast<const char*>(p.first), (std::streamsize)p.second);
  o.write(reinterpret_cast<const char*>(&l), 4);
  return (bool)o;
}

// reads one record into buf; returns its length or -1
int64_t get_record(std::ifstream& in, std::vector<char>& buf) {
  int32_t l = 0, t = 0;
  if (!in.read(reinterpret_cast<char*>(&l), 4) || l < 0) return -1;
  buf.resize((size_t)l);
  if (!in.read(buf.data(), l)) return -1;
  if (!in.read(reinterpret_cast<char*>(&t), 4) || t != l) return -1;
  return l;
}

}  // namespace

extern "C" {

int dgv_sol_file_name(const dgv_grid_desc* grid, const char* dir, const char* base, int32_t k_c, int32_t N_x,
                      const char* suff, char* out, size_t cap) {
  if (!grid || !out) { set_global_error("bad argument"); return DGV_ERR_ARG; }
  const std::string s = base_name(*grid, dir, base, k_c, N_x) + "_time" + trimmed(suff) + "_sol.dat";
  if (s.size() + 1 > cap) { set_global_error("output buffer too small"); return DGV_ERR_ARG; }
  std::memcpy(out, s.c_str(), s.size() + 1);
  return DGV_OK;
}

int dgv_log_file_name(const dgv_grid_desc* grid, const char* dir, const char* base, int32_t k_c, int32_t N_x,
                      const char* suffix, char* out, size_t cap) {
  if (!grid || !out) { set_global_error("bad argument"); return DGV_ERR_ARG; }
  const std::string s = base_name(*grid, dir, base, k_c, N_x) + (suffix ? suffix : "");
  if (s.size() + 1 > cap) { set_global_error("output buffer too small"); return DGV_ERR_ARG; }
  std::memcpy(out, s.c_str(), s.size() + 1);
  return DGV_OK;
}

int dgv_sol_write_file(const char* path, double e_time, double dt, int32_t rkmts, int64_t n, const double* f,
                       const double* hist) {
  if (!path || !f || n < 0 || n > INT32_MAX) { set_global_error("bad argument"); return DGV_ERR_ARG; }
  if (rkmts < 1 || rkmts > 5) {
    set_global_error("Unsupported degree of Multiple Time Stepping method passed");
    return DGV_ERR_ARG;
  }
  if (rkmts > 1 && !hist) { set_global_error("history arrays required for rkmts > 1"); return DGV_ERR_ARG; }
  std::ofstream o(path, std::ios::binary);
  if (!o) { set_global_error(std::string("cannot open ") + path); return DGV_ERR_IO; }
  const double td[2] = {e_time, dt};
  const int32_t m = (int32_t)n;
  const size_t nb = sizeof(double) * (size_t)n;
  std::vector<std::pair<const void*, size_t>> sol{{f, nb}};
  for (int k = 0; k < rkmts - 1; ++k) sol.push_back({hist + (size_t)k * n, nb});        // f1..f_{k-1}
  for (int k = 0; k < rkmts - 1; ++k) sol.push_back({hist + (size_t)(4 + k) * n, nb});  // frhs1..
  if (!put_record(o, {{td, 16}}) || !put_record(o, {{&rkmts, 4}}) || !put_record(o, {{&m, 4}}) ||
      !put_record(o, sol)) {
    set_global_error(std::string("write failed (or a record above 2 GiB): ") + path);
    return DGV_ERR_IO;
  }
  return DGV_OK;
}

int dgv_sol_write_init_file(const char* path, double e_time, double dt, int32_t rkmts, int64_t n,
                            const double* f) {
  if (!path || !f || n < 0 || n > INT32_MAX) { set_global_error("bad argument"); return DGV_ERR_ARG; }
  if (rkmts < 1 || rkmts > 5) {
    set_global_error("Unsupported degree of Multiple Time Stepping method passed");
    return DGV_ERR_ARG;
  }
  std::ofstream o(path, std::ios::binary);
  if (!o) { set_global_error(std::string("cannot open ") + path); return DGV_ERR_IO; }
  const double td[2] = {e_time, dt};
  const int32_t m = (int32_t)n;
  if (!put_record(o, {{td, 16}}) || !put_record(o, {{&rkmts, 4}}) || !put_record(o, {{&m, 4}}) ||
      !put_record(o, {{f, sizeof(double) * (size_t)n}})) {
    set_global_error(std::string("write failed (or a record above 2 GiB): ") + path);
    return DGV_ERR_IO;
  }
  return DGV_OK;
}

int dgv_sol_read_file(const char* path, double* e_time, double* dt, int32_t* rkmts, int64_t n, double* f,
                      double* hist) {
  if (!path || !f || n < 0) { set_global_error("bad argument"); return DGV_ERR_ARG; }
  std::ifstream in(path, std::ios::binary);
  if (!in) { set_global_error(std::string("cannot open ") + path); return DGV_ERR_IO; }
  std::vector<char> r;
  if (get_record(in, r) != 16) { set_global_error("solution file: bad time record"); return DGV_ERR_IO; }
  double td[2];
  std::memcpy(td, r.data(), 16);
  if (get_record(in, r) != 4) { set_global_error("solution file: bad rkmts record"); return DGV_ERR_IO; }
  int32_t k = 0;
  std::memcpy(&k, r.data(), 4);
  if (get_record(in, r) != 4) { set_global_error("solution file: bad size record"); return DGV_ERR_IO; }
  int32_t m = 0;
  std::memcpy(&m, r.data(), 4);
  if (m != n) {
    set_global_error("RestartSolutionDGV: trying to read a storage of wrong size, m/= size(f,1)");
    return DGV_ERR_ARG;
  }
  if (k < 1 || k > 5) {
    set_global_error("RestartSolutionDGV: Unsupported degree of Multiple Time Stepping method passed");
    return DGV_ERR_IO;
  }
  if (k > 1 && !hist) { set_global_error("history arrays required for rkmts > 1"); return DGV_ERR_ARG; }
  const size_t nb = sizeof(double) * (size_t)n;
  if (get_record(in, r) != (int64_t)((2 * k - 1) * nb)) {
    set_global_error("solution file: bad solution record");
    return DGV_ERR_IO;
  }
  std::memcpy(f, r.data(), nb);
  for (int j = 0; j < k - 1; ++j) std::memcpy(hist + (size_t)j * n, r.data() + (1 + j) * nb, nb);
  for (int j = 0; j < k - 1; ++j) std::memcpy(hist + (size_t)(4 + j) * n, r.data() + (k + j) * nb, nb);
  if (e_time) *e_time = td[0];
  if (dt) *dt = td[1];
  if (rkmts) *rkmts = k;
  return DGV_OK;
}

}  // extern "C"
