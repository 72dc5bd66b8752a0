// Wstat Davidson driver at Gamma on the device (FP64).
//
// Follows Wstat/davidson_diago.f90:31-437 (davidson_diago_gamma) with the real
// variants of Wstat/wstat_tools.f90 (build_hr 345-468, serial_diagox_dsy
// 208-273, redistribute_vr_distr 591-765, update_with_vr_distr 945-1092,
// refresh_with_vr_distr 1245-1395) on one process.  The basis dvg and its
// images dng stay resident on the device in stick order; the subspace
// products are matrix-core GEMMs; the dense eigensolve (matdiago_dsy,
// Tools/linear_algebra_kernel.f90:84-181) is rocSOLVER dsyevd in FP64.
//
// Orthonormalisation: do_mgs (davidson_diago.f90:909-1117) projects each old
// basis vector out of the new ones, one at a time, then runs Gram-Schmidt
// inside the new block.  Here the old-basis part is two classical
// Gram-Schmidt passes done as GEMMs (the same span, reorthogonalised), and
// the in-block part is the same sequential modified Gram-Schmidt.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>

#include "../../include/west_hip.h"
#include "wh_context.h"
#include "wh_gemm.h"
#include "wh_vec.h"

namespace wh {

namespace {

constexpr int NT = 256;
// do_randomize (davidson_diago.f90:1156-1177): vector ivec (1-based), plane
// wave ig (1-based, caller order): amplitude u0, phase 2 pi u1, divided by
// |G|^2 + 1 in (2 pi / alat)^2 units; G = 0 left at zero.
__global__ void k_randomize(int npw, int ivec0, int nvec, const int* perm, const double* g2kin, double inv_tpiba2,
                            z64* amat, long ld) {
  const int r = blockIdx.y;
  const int s = blockIdx.x * NT + threadIdx.x;
  if (s >= npw || r >= nvec) return;
  const unsigned long long ivec = static_cast<unsigned long long>(ivec0 + r);
  const unsigned long long ig = static_cast<unsigned long long>(perm[s]) + 1ULL;
  double u0, u1;
  start_uniforms(ivec, ig, u0, u1);
  z64 v = mk(0.0, 0.0);
  if (s != 0) {
    const double arg = 2.0 * 3.14159265358979323846 * u1;
    const double den = g2kin[s] * inv_tpiba2 + 1.0;
    v = mk(u0 * cos(arg) / den, u0 * sin(arg) / den);
  }
  amat[static_cast<long>(r) * ld + s] = v;
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < NT / 64; ++i) s += red[i];
  return s;
}

// v <- v / sqrt(<v|v>)   (davidson_diago.f90:1005-1025, Gamma norm)
__global__ __launch_bounds__(NT) void k_normalize(int npw, int gstart, z64* v) {
  __shared__ double red[NT / 64];
  double s = 0.0;
  for (int q = threadIdx.x; q < npw; q += NT) s += v[q].x * v[q].x + v[q].y * v[q].y;
  s = block_sum(s, red);
  double an = 2.0 * s;
  if (gstart == 2) an -= v[0].x * v[0].x;
  __syncthreads();
  const double f = 1.0 / sqrt(an);
  for (int q = threadIdx.x; q < npw; q += NT) v[q] = cscale(v[q], f);
}

__global__ void k_scale_rows(int m, int npw, const double* w, const z64* in, z64* out, long ld) {
  const int r = blockIdx.y;
  const int s = blockIdx.x * NT + threadIdx.x;
  if (s >= npw || r >= m) return;
  out[static_cast<long>(r) * ld + s] = cscale(in[static_cast<long>(r) * ld + s], w[s]);
}

struct Solver {
  rocblas_handle h = nullptr;
  ~Solver() {
    if (h) rocblas_destroy_handle(h);
  }
};

Solver& solver() {
  static thread_local Solver s;
  if (!s.h) {
    WH_REQUIRE(rocblas_create_handle(&s.h) == rocblas_status_success, "rocblas_create_handle failed");
  }
  WH_REQUIRE(rocblas_set_stream(s.h, ctx().stream) == rocblas_status_success, "rocblas_set_stream failed");
  return s;
}

// The subspace eigensolve (matdiago_dsy, Tools/linear_algebra_kernel.f90:84-181)
// on rocSOLVER: divide and conquer (dsyevd), QR iteration (dsyev) or Jacobi
// (dsyevj, tolerance at machine precision).  On matrices of a few hundred
// rows they differ several-fold in time; the first call in each size class
// (n rounded up to 32) times all three on copies of its matrix and keeps the
// fastest, and every call then runs that one.  Above kEigTimedMax rows
// dsyevd runs untimed: QR iteration alone takes seconds there (3.5 s at
// n = 1024), more than the timing could ever win back.
constexpr int kEigTimedMax = 256;
std::mutex g_eig_mutex;
std::map<int, int>& eig_choice() {
  static std::map<int, int> m;
  return m;
}
const char* const kEigName[3] = {"dsyevd", "dsyev", "dsyevj"};

void run_syev(rocblas_handle h, int v, int n, double* A, double* D, double* E, rocblas_int* info, double* resid,
              rocblas_int* nsw) {
  rocblas_status st;
  if (v == 0) st = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, n, A, n, D, E, info);
  else if (v == 1) st = rocsolver_dsyev(h, rocblas_evect_original, rocblas_fill_upper, n, A, n, D, E, info);
  else st = rocsolver_dsyevj(h, rocblas_esort_ascending, rocblas_evect_original, rocblas_fill_upper, n, A, n, 0.0, resid,
                             100, nsw, D, info);
  WH_REQUIRE(st == rocblas_status_success, std::string("rocsolver_") + kEigName[v] + " failed");
}

// glbrak of rows: C (na x nb, col major ldc) = <A_i|B_j>
void glbrak_rows(const z64* A, int na, const z64* B, int nb, double* Cm, int ldc) {
  Context& c = ctx();
  const long k2 = 2L * c.npw;
  const double* ar = reinterpret_cast<const double*>(A);
  const double* br = reinterpret_cast<const double*>(B);
  WH_TIMED("gemm_bra", dgemm_bra(na, nb, static_cast<int>(k2), 2.0, ar, k2, br, k2, c.gstart == 2 ? 1.0 : 0.0,
                                 nullptr, Cm, ldc, c.stream));
}

// Y (nb rows) = alpha * X (na rows)^T-combination: Y_j = alpha sum_i W(i,j) X_i + beta Y_j
void combine_rows(const z64* X, int na, const double* W, int ldw, z64* Y, int nb, double alpha, double beta) {
  Context& c = ctx();
  const long k2 = 2L * c.npw;
  WH_TIMED("gemm_ket", dgemm(false, static_cast<int>(k2), nb, na, alpha, reinterpret_cast<const double*>(X), k2, W,
                             ldw, beta, reinterpret_cast<double*>(Y), k2, c.stream));
}

// orthonormalise rows [ms, me) of amat against [0, ms) and among themselves
void do_mgs(z64* amat, int ms, int me) {
  Context& c = ctx();
  hipStream_t s = c.stream;
  const int npw = c.npw;
  const int nnew = me - ms;
  if (nnew <= 0) return;
  if (ms > 0) {
    double* cm = c.w_cmat.as<double>(static_cast<size_t>(ms) * nnew);
    for (int pass = 0; pass < 2; ++pass) {
      glbrak_rows(amat, ms, amat + static_cast<long>(ms) * npw, nnew, cm, ms);
      combine_rows(amat, ms, cm, ms, amat + static_cast<long>(ms) * npw, nnew, -1.0, 1.0);
    }
  }
  double* cm = c.w_cmat.as<double>(static_cast<size_t>(std::max(nnew, 1)));
  for (int k = ms; k < me; ++k) {
    z64* vk = amat + static_cast<long>(k) * npw;
    k_normalize<<<1, NT, 0, s>>>(npw, c.gstart, vk);
    WH_HIP(hipGetLastError());
    const int nr = me - k - 1;
    if (nr <= 0) continue;
    glbrak_rows(vk, 1, vk + npw, nr, cm, 1);
    combine_rows(vk, 1, cm, 1, vk + npw, nr, -1.0, 1.0);
  }
}

// ---- restart files (Wstat/davidson_restart.f90) ---------------------------

constexpr int kHdLength = 32;          // IO_kernel/mod_west_io.f90:31-35
constexpr int kHdVersion = 210405;

bool little_endian() {
  const uint32_t one = 1;
  unsigned char b;
  std::memcpy(&b, &one, 1);
  return b == 1;
}

std::string vec_name(const std::string& dir, char kind, int j) {   // V/N + i6.6 (davidson_restart.f90:151-203)
  char buf[16];
  std::snprintf(buf, sizeof buf, "%c%06d.dat", kind, j);
  return dir + "/" + buf;
}

// one basis vector: header, then n complex coefficients (pdep_merge_and_write_G,
// IO_kernel/pdep_io.f90:115-141)
void write_vec(const std::string& fname, const z64* v, int n) {
  int32_t hd[kHdLength] = {};
  hd[0] = kHdVersion;
  hd[1] = little_endian() ? 1 : 0;
  hd[2] = n;
  FILE* f = std::fopen(fname.c_str(), "wb");
  WH_REQUIRE(f != nullptr, "davidson_restart_write: cannot open file: " + fname);
  const bool ok = std::fwrite(hd, sizeof hd, 1, f) == 1 && std::fwrite(v, sizeof(z64), n, f) == static_cast<size_t>(n);
  std::fclose(f);
  WH_REQUIRE(ok, "davidson_restart_write: cannot write file: " + fname);
}

// pdep_read_G_and_distribute (pdep_io.f90:254-285), with its checks
void read_vec(const std::string& fname, z64* v, int n) {
  FILE* f = std::fopen(fname.c_str(), "rb");
  WH_REQUIRE(f != nullptr, "pdep_read: Cannot read file: " + fname);
  int32_t hd[kHdLength];
  const bool got = std::fread(hd, sizeof hd, 1, f) == 1;
  const bool ok_version = got && hd[0] == kHdVersion;
  const bool ok_dim = ok_version && hd[2] == n;
  const bool ok_endian = ok_dim && (hd[1] == 1) == little_endian();
  const bool ok_data = ok_endian && std::fread(v, sizeof(z64), n, f) == static_cast<size_t>(n);
  std::fclose(f);
  WH_REQUIRE(ok_version, "pdep_read: Unknown file format: " + fname);
  WH_REQUIRE(ok_dim, "pdep_read: Dimension mismatch: " + fname);
  WH_REQUIRE(ok_endian, "pdep_read: Endianness mismatch: " + fname);
  WH_REQUIRE(ok_data, "pdep_read: Cannot read file: " + fname);
}

// hr_vr.dat: two Fortran unformatted sequential records, hr then vr
// (davidson_restart.f90:117-130; 4-byte length markers around each record)
void write_hr_vr(const std::string& dir, const std::vector<double>& hr, const std::vector<double>& vr) {
  const std::string fname = dir + "/hr_vr.dat";
  const size_t bytes = hr.size() * sizeof(double);
  WH_REQUIRE(bytes < (size_t(1) << 31), "davidson_restart_write: hr too large for one record");
  const int32_t mark = static_cast<int32_t>(bytes);
  FILE* f = std::fopen(fname.c_str(), "wb");
  WH_REQUIRE(f != nullptr, "davidson_restart_write: cannot open file: " + fname);
  bool ok = true;
  for (const std::vector<double>* a : {&hr, &vr})
    ok = ok && std::fwrite(&mark, 4, 1, f) == 1 && std::fwrite(a->data(), 1, bytes, f) == bytes &&
         std::fwrite(&mark, 4, 1, f) == 1;
  std::fclose(f);
  WH_REQUIRE(ok, "davidson_restart_write: cannot write file: " + fname);
}

void read_hr_vr(const std::string& dir, std::vector<double>& hr, std::vector<double>& vr) {
  const std::string fname = dir + "/hr_vr.dat";
  FILE* f = std::fopen(fname.c_str(), "rb");
  WH_REQUIRE(f != nullptr, "davidson_restart_read: cannot read file: " + fname);
  const size_t bytes = hr.size() * sizeof(double);
  bool ok = true;
  for (std::vector<double>* a : {&hr, &vr}) {
    int32_t m0 = 0, m1 = 0;
    ok = ok && std::fread(&m0, 4, 1, f) == 1 && static_cast<size_t>(m0) == bytes &&
         std::fread(a->data(), 1, bytes, f) == bytes && std::fread(&m1, 4, 1, f) == 1 && m1 == m0;
  }
  std::fclose(f);
  WH_REQUIRE(ok, "davidson_restart_read: hr_vr.dat does not hold two records of n_pdep_basis^2 reals: " + fname);
}

// summary.json (davidson_restart.f90:104-115): dav_iter, notcnv, nbase, conv, ev, ew
void write_summary(const std::string& dir, int dav_iter, int notcnv, int nbase, const std::vector<char>& conv,
                   const std::vector<double>& ev, const std::vector<double>& ew) {
  const std::string fname = dir + "/summary.json";
  FILE* f = std::fopen(fname.c_str(), "w");
  WH_REQUIRE(f != nullptr, "davidson_restart_write: cannot open file: " + fname);
  std::fprintf(f, "{\n  \"dav_iter\": %d,\n  \"notcnv\": %d,\n  \"nbase\": %d,\n  \"conv\": [", dav_iter, notcnv, nbase);
  for (size_t i = 0; i < conv.size(); ++i) std::fprintf(f, "%s%s", i ? ", " : "", conv[i] ? "true" : "false");
  std::fprintf(f, "],\n  \"ev\": [");
  for (size_t i = 0; i < ev.size(); ++i) std::fprintf(f, "%s%.17g", i ? ", " : "", ev[i]);
  std::fprintf(f, "],\n  \"ew\": [");
  for (size_t i = 0; i < ew.size(); ++i) std::fprintf(f, "%s%.17g", i ? ", " : "", ew[i]);
  const bool ok = std::fprintf(f, "]\n}\n") > 0;
  std::fclose(f);
  WH_REQUIRE(ok, "davidson_restart_write: cannot write file: " + fname);
}

// the value text after "key": in a flat JSON object (the summary's layout)
const char* json_value(const std::string& doc, const char* key) {
  const std::string k = std::string("\"") + key + "\"";
  const size_t at = doc.find(k);
  if (at == std::string::npos) return nullptr;
  const size_t colon = doc.find(':', at + k.size());
  return colon == std::string::npos ? nullptr : doc.c_str() + colon + 1;
}

bool json_int(const std::string& doc, const char* key, int& out) {
  const char* p = json_value(doc, key);
  if (!p) return false;
  char* end = nullptr;
  const long v = std::strtol(p, &end, 10);
  if (end == p) return false;
  out = static_cast<int>(v);
  return true;
}

// array of numbers or of true/false
bool json_array(const std::string& doc, const char* key, std::vector<double>& out) {
  const char* p = json_value(doc, key);
  if (!p) return false;
  while (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t') ++p;
  if (*p != '[') return false;
  ++p;
  out.clear();
  for (;;) {
    while (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t' || *p == ',') ++p;
    if (*p == ']') return true;
    if (std::strncmp(p, "true", 4) == 0) {
      out.push_back(1.0);
      p += 4;
    } else if (std::strncmp(p, "false", 5) == 0) {
      out.push_back(0.0);
      p += 5;
    } else {
      char* end = nullptr;
      const double v = std::strtod(p, &end);
      if (end == p) return false;
      out.push_back(v);
      p = end;
    }
  }
}

void remove_restart(const std::string& dir, int n_pdep_basis) {   // davidson_restart_clear (360-410)
  std::remove((dir + "/summary.json").c_str());
  std::remove((dir + "/hr_vr.dat").c_str());
  for (int j = 1; j <= n_pdep_basis; ++j) {
    std::remove(vec_name(dir, 'V', j).c_str());
    std::remove(vec_name(dir, 'N', j).c_str());
  }
  ::rmdir(dir.c_str());
}

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

}  // namespace wh

using namespace wh;

extern "C" int wh_davidson_diago(const wh_davidson_params* prm, double* ev_out, int* conv_out, double* dvg_out,
                                 wh_davidson_step* log, int log_capacity, int* log_len, double* ev_log,
                                 int* conv_log) {
  try {
    Context& c = ctx();
    require_ready(true);
    WH_REQUIRE(prm != nullptr, "params required");
    WH_REQUIRE(c.have_sqvc, "wh_set_coulomb has not been called");
    hipStream_t s = c.stream;
    const int nvec = prm->n_pdep_eigen;
    const int nvecx = prm->n_pdep_eigen * prm->n_pdep_times;
    WH_REQUIRE(nvec > 0, "n_pdep_eigen must be positive");
    WH_REQUIRE(nvec <= nvecx / 2, "nvecx is too small");   // davidson_diago.f90:108
    const int npw = c.npw;
    const long ld = npw;
    if (log_len) *log_len = 0;
    // device basis
    DevBuf &b_dvg = c.dav[0], &b_dng = c.dav[1], &b_tmp = c.dav[2], &b_hr = c.dav[3], &b_w = c.dav[4],
           &b_e = c.dav[5], &b_info = c.dav[6], &b_vr = c.dav[7], &b_sq = c.dav[8];
    (void)b_sq;
    z64* dvg = b_dvg.as<z64>(static_cast<size_t>(nvecx) * npw);
    z64* dng = b_dng.as<z64>(static_cast<size_t>(nvecx) * npw);
    z64* tmp = b_tmp.as<z64>(static_cast<size_t>(nvecx) * npw);
    WH_HIP(hipMemsetAsync(dvg, 0, sizeof(z64) * nvecx * npw, s));
    WH_HIP(hipMemsetAsync(dng, 0, sizeof(z64) * nvecx * npw, s));
    std::vector<double> hr(static_cast<size_t>(nvecx) * nvecx, 0.0), vr(static_cast<size_t>(nvecx) * nvecx, 0.0);
    std::vector<double> ew(nvecx, 0.0), ev(nvec, 0.0);
    std::vector<char> conv(nvec, 0);
    int nbase = nvec, notcnv = nvec;
    const double t0 = now_s();
    double tit = t0;
    long ns0 = c.nstern_calls;
    int nlog = 0;

    // chi~ applied to rows [r0, r0+n): dng = sqvc chi0 (sqvc dvg)
    auto apply_op = [&](int r0, int n, double tr2) {
      z64* in = tmp;   // sqvc * dvg rows
      k_scale_rows<<<dim3(ceil_div(npw, NT), n), NT, 0, s>>>(n, npw, c.d_sqvc, dvg + static_cast<long>(r0) * ld, in, ld);
      WH_HIP(hipGetLastError());
      dfpt_dev(n, in, dng + static_cast<long>(r0) * ld, tr2, c.d_sqvc, nullptr);
    };
    // hr(0:g_e, cols) = <dvg_i|dng_j>, rows beyond g_e zeroed
    auto build_hr = [&](int l2s, int l2e, int g_e) {
      const int nb = l2e - l2s;
      double* d = b_hr.as<double>(static_cast<size_t>(g_e) * nb);
      glbrak_rows(dvg, g_e, dng + static_cast<long>(l2s) * ld, nb, d, g_e);
      std::vector<double> blk(static_cast<size_t>(g_e) * nb);
      WH_HIP(hipMemcpyAsync(blk.data(), d, sizeof(double) * blk.size(), hipMemcpyDeviceToHost, s));
      WH_HIP(hipStreamSynchronize(s));
      for (int j = 0; j < nb; ++j) {
        double* col = &hr[static_cast<size_t>(l2s + j) * nvecx];
        std::fill(col, col + nvecx, 0.0);
        std::memcpy(col, &blk[static_cast<size_t>(j) * g_e], sizeof(double) * g_e);
      }
    };
    // serial_diagox_dsy: eigenpairs of hr(0:n,0:n) upper; nvec lowest kept
    auto diagox = [&](int n) {
      double* A = b_hr.as<double>(static_cast<size_t>(n) * n);
      double* D = b_w.as<double>(n);
      double* E = b_e.as<double>(n);
      rocblas_int* info = b_info.as<rocblas_int>(4);
      double* resid = reinterpret_cast<double*>(info + 2);
      rocblas_int* nsw = info + 1;
      std::vector<double> a(static_cast<size_t>(n) * n);
      for (int j = 0; j < n; ++j) std::memcpy(&a[static_cast<size_t>(j) * n], &hr[static_cast<size_t>(j) * nvecx], sizeof(double) * n);
      rocblas_handle h = solver().h;
      int pick = 0;
      {
        std::lock_guard<std::mutex> lk(g_eig_mutex);
        const int cls = (n + 31) / 32;
        auto it = eig_choice().find(cls);
        if (it == eig_choice().end() && n > kEigTimedMax) it = eig_choice().emplace(cls, 0).first;
        if (it == eig_choice().end()) {
          // time the three on copies (each run once untimed first: its workspace)
          float t[3] = {0.f, 0.f, 0.f};
          hipEvent_t e0, e1;
          WH_HIP(hipEventCreate(&e0));
          WH_HIP(hipEventCreate(&e1));
          for (int v = 0; v < 3; ++v)
            for (int rep = 0; rep < 2; ++rep) {
              WH_HIP(hipMemcpyAsync(A, a.data(), sizeof(double) * a.size(), hipMemcpyHostToDevice, s));
              WH_HIP(hipEventRecord(e0, s));
              run_syev(h, v, n, A, D, E, info, resid, nsw);
              WH_HIP(hipEventRecord(e1, s));
              WH_HIP(hipEventSynchronize(e1));
              WH_HIP(hipEventElapsedTime(&t[v], e0, e1));
            }
          WH_HIP(hipEventDestroy(e0));
          WH_HIP(hipEventDestroy(e1));
          for (int v = 1; v < 3; ++v)
            if (t[v] < t[pick]) pick = v;
          if (tune_log())
            std::fprintf(stderr, "[west_hip] subspace eigensolve n=%d: dsyevd %.3f dsyev %.3f dsyevj %.3f ms -> %s\n",
                         n, t[0], t[1], t[2], kEigName[pick]);
          it = eig_choice().emplace(cls, pick).first;
        }
        pick = it->second;
      }
      WH_HIP(hipMemcpyAsync(A, a.data(), sizeof(double) * a.size(), hipMemcpyHostToDevice, s));
      run_syev(h, pick, n, A, D, E, info, resid, nsw);
      std::vector<double> w(n);
      rocblas_int hinfo = 0;
      WH_HIP(hipMemcpyAsync(a.data(), A, sizeof(double) * a.size(), hipMemcpyDeviceToHost, s));
      WH_HIP(hipMemcpyAsync(w.data(), D, sizeof(double) * n, hipMemcpyDeviceToHost, s));
      WH_HIP(hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, s));
      WH_HIP(hipStreamSynchronize(s));
      WH_REQUIRE(hinfo == 0, std::string(kEigName[pick]) + " did not converge");
      for (int j = 0; j < nvec; ++j) {
        ew[j] = w[j];
        std::memcpy(&vr[static_cast<size_t>(j) * nvecx], &a[static_cast<size_t>(j) * n], sizeof(double) * n);
      }
    };
    auto record = [&](int dav_iter, double tr2, int dfpt_dim, int diago_dim) {
      if (log && nlog < log_capacity) {
        wh_davidson_step& r = log[nlog];
        r.dav_iter = dav_iter;
        r.notcnv = notcnv;
        r.dfpt_dim = dfpt_dim;
        r.diago_dim = diago_dim;
        r.sternop_ncalls = c.nstern_calls - ns0;
        r.dfpt_tr2 = tr2;
        r.time_elap = now_s() - t0;
        r.time_iter = now_s() - tit;
      }
      // ev and conv of the step (output_ev_and_time, davidson_diago.f90:1318-1319)
      if (nlog < log_capacity) {
        if (ev_log) std::memcpy(ev_log + static_cast<size_t>(nlog) * nvec, ev.data(), sizeof(double) * nvec);
        if (conv_log)
          for (int n = 0; n < nvec; ++n) conv_log[static_cast<size_t>(nlog) * nvec + n] = conv[n] ? 1 : 0;
      }
      ++nlog;
      // the per-step report of the reference (output_ev_and_time, davidson_diago.f90:355)
      if (tune_log())
        std::fprintf(stderr, "[west_hip] davidson iter %d: notcnv %d, dfpt_dim %d, diago_dim %d, %.2f s\n",
                     dav_iter, notcnv, dfpt_dim, diago_dim, now_s() - t0);
    };
    // rows [0, ncol) of X <- X * VR(0:n, 0:ncol) (refresh); uses tmp
    auto refresh = [&](z64* X, int n, int ncol) {
      double* W = b_vr.as<double>(static_cast<size_t>(n) * ncol);
      std::vector<double> w(static_cast<size_t>(n) * ncol);
      for (int j = 0; j < ncol; ++j) std::memcpy(&w[static_cast<size_t>(j) * n], &vr[static_cast<size_t>(j) * nvecx], sizeof(double) * n);
      WH_HIP(hipMemcpyAsync(W, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice, s));
      combine_rows(X, n, W, n, tmp, ncol, 1.0, 0.0);
      WH_HIP(hipMemcpyAsync(X, tmp, sizeof(z64) * ncol * npw, hipMemcpyDeviceToDevice, s));
      if (nvecx > ncol) WH_HIP(hipMemsetAsync(X + static_cast<long>(ncol) * ld, 0, sizeof(z64) * (nvecx - ncol) * npw, s));
      WH_HIP(hipStreamSynchronize(s));
    };

    // ---- restart files (Wstat/davidson_restart.f90): the state after a step,
    // basis rows in the caller's G order
    const std::string rdir = c.restart_dir;
    const int nsw = c.n_steps_write_restart;
    auto restart_write = [&](int dav_iter_w) {   // davidson_restart_write_real (33-222)
      ::mkdir(rdir.c_str(), 0755);
      write_summary(rdir, dav_iter_w, notcnv, nbase, conv, ev, ew);
      write_hr_vr(rdir, hr, vr);
      const size_t cin = static_cast<size_t>(nbase) * c.npwx;
      std::vector<z64> h(cin);
      z64* o = c.w_out.as<z64>(cin);
      for (int k = 0; k < 2; ++k) {
        scatter_caller(nbase, npw, c.npwx, c.d_perm, k ? dng : dvg, ld, o, c.npwx, s);
        WH_HIP(hipMemcpyAsync(h.data(), o, sizeof(z64) * cin, hipMemcpyDeviceToHost, s));
        WH_HIP(hipStreamSynchronize(s));
        for (int j = 0; j < nbase; ++j)
          write_vec(vec_name(rdir, k ? 'N' : 'V', j + 1), h.data() + static_cast<size_t>(j) * c.npwx, npw);
      }
    };
    auto restart_read = [&]() {   // davidson_restart_read_real (413-462)
      std::ifstream in(rdir + "/summary.json");
      WH_REQUIRE(in.good(), "davidson_restart_read: cannot read file: " + rdir + "/summary.json");
      std::stringstream ss;
      ss << in.rdbuf();
      const std::string doc = ss.str();
      int di = 0;
      std::vector<double> a_conv, a_ev, a_ew;
      WH_REQUIRE(json_int(doc, "dav_iter", di) && json_int(doc, "notcnv", notcnv) && json_int(doc, "nbase", nbase) &&
                     json_array(doc, "conv", a_conv) && json_array(doc, "ev", a_ev) && json_array(doc, "ew", a_ew),
                 "davidson_restart_read: summary.json lacks dav_iter, notcnv, nbase, conv, ev or ew");
      WH_REQUIRE(static_cast<int>(a_conv.size()) >= nvec && static_cast<int>(a_ev.size()) == nvec &&
                     static_cast<int>(a_ew.size()) == nvecx && nbase >= nvec && nbase <= nvecx && notcnv >= 0 &&
                     notcnv <= nvec && di >= -1,
                 "davidson_restart_read: the restart does not match n_pdep_eigen and n_pdep_times");
      for (int n = 0; n < nvec; ++n) {
        conv[n] = a_conv[n] != 0.0;
        ev[n] = a_ev[n];
      }
      for (int n = 0; n < nvecx; ++n) ew[n] = a_ew[n];
      read_hr_vr(rdir, hr, vr);
      const size_t cin = static_cast<size_t>(nbase) * c.npwx;
      std::vector<z64> h(cin, z64{0.0, 0.0});
      z64* o = c.w_out.as<z64>(cin);
      for (int k = 0; k < 2; ++k) {
        for (int j = 0; j < nbase; ++j)
          read_vec(vec_name(rdir, k ? 'N' : 'V', j + 1), h.data() + static_cast<size_t>(j) * c.npwx, npw);
        WH_HIP(hipMemcpyAsync(o, h.data(), sizeof(z64) * cin, hipMemcpyHostToDevice, s));
        gather_sticks(nbase, npw, c.d_perm, o, c.npwx, k ? dng : dvg, ld, nullptr, s);
        WH_HIP(hipStreamSynchronize(s));
      }
      return di;
    };

    int dav_iter = -1;
    if (c.restart_from) {
      dav_iter = restart_read();   // (154-159)
      c.restart_from = 0;          // later runs start from scratch (as after a restart, 613)
    } else {
      // ---- from scratch (davidson_diago.f90:167-196)
      k_randomize<<<dim3(ceil_div(npw, NT), nvec), NT, 0, s>>>(npw, 1, nvec, c.d_perm, c.d_g2kin, 1.0 / c.tpiba2,
                                                                dvg, ld);
      WH_HIP(hipGetLastError());
      do_mgs(dvg, 0, nvec);
      apply_op(0, nvec, -1.0);
      if (nsw == 1) restart_write(dav_iter);
    }
    double tr2 = std::min(1.0e-4, 1.0e8 * prm->tr2_dfpt);
    if (dav_iter == -1) {
      // ---- extra step (205-249)
      tit = now_s();
      ns0 = c.nstern_calls;
      WH_HIP(hipMemcpyAsync(dvg, dng, sizeof(z64) * nvec * npw, hipMemcpyDeviceToDevice, s));
      do_mgs(dvg, 0, nvec);
      apply_op(0, nvec, tr2);
      build_hr(0, nvec, nvec);
      diagox(nbase);
      for (int i = 0; i < nvec; ++i) ev[i] = ew[i];
      record(0, tr2, nbase, nbase);
      dav_iter = 0;
      if (nsw == 1) restart_write(dav_iter);
    }
    bool converged = false;
    for (int kter = 1; kter <= prm->n_pdep_maxiter; ++kter) {   // (259-424)
      tit = now_s();
      ns0 = c.nstern_calls;
      ++dav_iter;
      const int dfpt_dim = notcnv;
      const int diago_dim = nbase + notcnv;
      int np = 0;
      for (int n = 0; n < nvec; ++n) {
        if (!conv[n]) {
          std::memcpy(&vr[static_cast<size_t>(nbase + np) * nvecx], &vr[static_cast<size_t>(n) * nvecx], sizeof(double) * nvecx);
          ew[nbase + np] = ev[n];
          ++np;
        }
      }
      // update_with_vr_distr: new_j = sum_i vr(i,j) (dng_i - ew_j dvg_i)
      {
        double* W = b_vr.as<double>(static_cast<size_t>(nbase) * 2 * notcnv);
        std::vector<double> w(static_cast<size_t>(nbase) * 2 * notcnv);
        for (int j = 0; j < notcnv; ++j)
          for (int i = 0; i < nbase; ++i) {
            const double v = vr[static_cast<size_t>(nbase + j) * nvecx + i];
            w[static_cast<size_t>(j) * nbase + i] = v;
            w[static_cast<size_t>(notcnv + j) * nbase + i] = -ew[nbase + j] * v;
          }
        WH_HIP(hipMemcpyAsync(W, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice, s));
        z64* newv = dvg + static_cast<long>(nbase) * ld;
        combine_rows(dng, nbase, W, nbase, newv, notcnv, 1.0, 0.0);
        combine_rows(dvg, nbase, W + static_cast<size_t>(notcnv) * nbase, nbase, newv, notcnv, 1.0, 1.0);
        WH_HIP(hipStreamSynchronize(s));
      }
      do_mgs(dvg, nbase, nbase + notcnv);
      tr2 = prm->tr2_dfpt;
      apply_op(nbase, notcnv, tr2);
      build_hr(nbase, nbase + notcnv, nbase + notcnv);
      nbase += notcnv;
      diagox(nbase);
      notcnv = 0;
      for (int n = 0; n < nvec; ++n) {
        conv[n] = (std::fabs((ew[n] - ev[n]) / ev[n]) < prm->trev_pdep_rel) && (std::fabs(ew[n] - ev[n]) < prm->trev_pdep);
        if (!conv[n]) ++notcnv;
      }
      for (int n = 0; n < nvec; ++n) ev[n] = ew[n];
      record(dav_iter, tr2, dfpt_dim, diago_dim);
      if (notcnv == 0 || nbase + notcnv > nvecx || kter == prm->n_pdep_maxiter) {
        if (notcnv == 0) {
          refresh(dvg, nbase, nvec);
          converged = true;
          break;
        }
        if (kter == prm->n_pdep_maxiter) break;
        refresh(dvg, nbase, nvec);
        refresh(dng, nbase, nvec);
        nbase = nvec;
        std::fill(hr.begin(), hr.end(), 0.0);
        std::fill(vr.begin(), vr.end(), 0.0);
        for (int i = 0; i < nbase; ++i) {
          hr[static_cast<size_t>(i) * nvecx + i] = ev[i];
          vr[static_cast<size_t>(i) * nvecx + i] = 1.0;
        }
      }
      if (nsw > 0 && dav_iter % nsw == 0) restart_write(dav_iter);   // (421-422)
    }
    if (converged && !rdir.empty()) remove_restart(rdir, nvecx);   // davidson_restart_clear (376)
    std::memcpy(ev_out, ev.data(), sizeof(double) * nvec);
    if (conv_out)
      for (int n = 0; n < nvec; ++n) conv_out[n] = conv[n] ? 1 : 0;
    if (dvg_out) {
      const size_t cin = static_cast<size_t>(nvec) * c.npwx;
      std::vector<z64> h(cin);
      z64* o = c.w_out.as<z64>(cin);
      scatter_caller(nvec, npw, c.npwx, c.d_perm, dvg, ld, o, c.npwx, s);
      WH_HIP(hipMemcpyAsync(h.data(), o, sizeof(z64) * cin, hipMemcpyDeviceToHost, s));
      WH_HIP(hipStreamSynchronize(s));
      WH_HIP(hipMemcpy(dvg_out, h.data(), sizeof(z64) * cin, hipMemcpyDefault));
    }
    WH_HIP(hipStreamSynchronize(s));
    if (c.prof.on) c.prof.flush();
    if (log_len) *log_len = std::min(nlog, log_capacity);
    return notcnv;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code < 0 ? e.code : -1;
  } catch (const std::exception& e) {
    set_error(e.what());
    return -3;
  }
}
