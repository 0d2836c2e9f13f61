// skinny_svqb: how the precision of the Gram matrix limits SVQB
// orthogonalisation (the experiment of svqb/svqb.cu:241-335).
//
//   skinny_svqb [--K 10000000] [--M 3] [--reps 3] [--min-pert 1e-6]
//
// W (K x M, row-major) holds 1 + pert * U[0,1) with two large entries
// (+-1e5 at flat indices 234 and 235), so its columns become nearly
// collinear as pert shrinks.  For each Gram variant (FP32, FP64,
// double-double, Ozaki with 3 limbs) one SVQB step runs: S = W^T W;
// D = diag(S)^-1/2; the eigen-decomposition D S D = U L U^T (cyclic Jacobi on
// the host, M is small); T = D U L^-1/2; Q = W T (TSMM).  The table reports
// max |Q^T Q - I| with Q^T Q formed in double-double, averaged over --reps
// draws, one row per pert = 4, 2, 1, ... >= --min-pert.  The last column is
// the library's own device SVQB (FP64 Gram, on-device Jacobi, in place).
#include <cmath>
#include <iomanip>
#include <iostream>

#include "common.hpp"

namespace {

// Cyclic Jacobi eigen-decomposition of the symmetric n x n matrix a
// (column-major, destroyed); eigenvalues in w, eigenvectors in the columns
// of v.
void jacobi(int n, std::vector<double>& a, std::vector<double>& w,
            std::vector<double>& v) {
  v.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) v[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0, dia = 0;
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++)
        (i == j ? dia : off) += a[(size_t)j * n + i] * a[(size_t)j * n + i];
    if (off <= 1e-30 * dia) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = a[(size_t)q * n + p];
        if (apq == 0.0) continue;
        const double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) /
                         (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {  // columns p, q
          const double akp = a[(size_t)p * n + k], akq = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = c * akp - s * akq;
          a[(size_t)q * n + k] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {  // rows p, q
          const double apk = a[(size_t)k * n + p], aqk = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = c * apk - s * aqk;
          a[(size_t)k * n + q] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          const double vkp = v[(size_t)p * n + k], vkq = v[(size_t)q * n + k];
          v[(size_t)p * n + k] = c * vkp - s * vkq;
          v[(size_t)q * n + k] = s * vkp + c * vkq;
        }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; i++) w[i] = a[(size_t)i * n + i];
}

// T (column-major M x M) from the Gram matrix S (column-major)
std::vector<double> svqb_transform(int M, std::vector<double> S) {
  std::vector<double> d(M);
  for (int i = 0; i < M; i++) {
    const double sii = S[(size_t)i * M + i];
    d[i] = sii > 1e-300 ? 1.0 / std::sqrt(sii) : 0.0;
  }
  for (int j = 0; j < M; j++)
    for (int i = 0; i < M; i++) S[(size_t)j * M + i] *= d[i] * d[j];
  std::vector<double> w, U;
  jacobi(M, S, w, U);
  std::vector<double> T((size_t)M * M);
  for (int j = 0; j < M; j++) {
    const double sc = 1.0 / std::sqrt(std::max(1e-15, w[j]));
    for (int i = 0; i < M; i++) T[(size_t)j * M + i] = d[i] * U[(size_t)j * M + i] * sc;
  }
  return T;
}

}  // namespace

int main(int argc, char** argv) {
  int64_t K = 10000000;
  int M = 3, reps = 3;
  double min_pert = 1e-6;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    if (a == "--K" && i + 1 < argc) K = std::stoll(argv[++i]);
    else if (a == "--M" && i + 1 < argc) M = std::stoi(argv[++i]);
    else if (a == "--reps" && i + 1 < argc) reps = std::stoi(argv[++i]);
    else if (a == "--min-pert" && i + 1 < argc) min_pert = std::stod(argv[++i]);
    else {
      std::fprintf(stderr, "usage: skinny_svqb [--K N] [--M N] [--reps N] [--min-pert X]\n");
      return 2;
    }
  }
  if (M < 1 || M > 64 || K < 300) {
    std::fprintf(stderr, "need 1 <= M <= 64 and K >= 300\n");
    return 2;
  }
  const size_t n = (size_t)K * M;
  drv::DevBuf<double> W(n), Q(n), S(M * M), T(M * M);
  drv::DevBuf<float> Wf(n), Sf(M * M);
  std::vector<double> hW(n), hS(M * M), hQQ(M * M);
  std::vector<float> hWf(n), hSf(M * M);
  skinny::set_abort_on_error(true);

  // Gram of W in the given precision, to the host (column-major)
  enum { FP32, FP64, DD, OZ3, DEVICE, NV };
  const char* names[NV] = {"FP32 Gram", "FP64 Gram", "dd Gram", "Ozaki-3 Gram",
                           "device SVQB"};
  auto gram = [&](int v) {
    if (v == FP32) {
      SKINNY_CHECK(skinny_stsmttsm(nullptr, 0, M, M, K, Wf.p, M, 1.f, Wf.p, M, 0.f,
                                   Sf.p, M, nullptr));
      HIP_CHECK(hipMemcpy(hSf.data(), Sf.p, sizeof(float) * M * M, hipMemcpyDeviceToHost));
      for (int i = 0; i < M * M; i++) hS[i] = hSf[i];
      return;
    }
    if (v == FP64)
      SKINNY_CHECK(skinny_dtsmttsm(nullptr, 0, M, M, K, W.p, M, 1.0, W.p, M, 0.0, S.p, M,
                                   nullptr));
    else if (v == DD)
      SKINNY_CHECK(skinny_htsmttsm(nullptr, 0, M, M, K, W.p, M, 1.0, W.p, M, 0.0, S.p, M,
                                   nullptr));
    else
      SKINNY_CHECK(skinny_dtsmttsm_ozaki(nullptr, 0, M, M, K, W.p, M, 1.0, W.p, M, 0.0,
                                         S.p, M, 3, nullptr));
    HIP_CHECK(hipMemcpy(hS.data(), S.p, sizeof(double) * M * M, hipMemcpyDeviceToHost));
  };
  // max |Q^T Q - I| with the Gram in double-double
  auto max_error = [&](const double* q) {
    SKINNY_CHECK(skinny_htsmttsm(nullptr, 0, M, M, K, q, M, 1.0, q, M, 0.0, S.p, M,
                                 nullptr));
    HIP_CHECK(hipMemcpy(hQQ.data(), S.p, sizeof(double) * M * M, hipMemcpyDeviceToHost));
    double e = 0;
    for (int j = 0; j < M; j++)
      for (int i = 0; i < M; i++)
        e = std::max(e, std::fabs(hQQ[(size_t)j * M + i] - (i == j ? 1.0 : 0.0)));
    return e;
  };

  std::cout << "| pert |";
  for (int v = 0; v < NV; v++) std::cout << " " << names[v] << " |";
  std::cout << "\n|---|";
  for (int v = 0; v < NV; v++) std::cout << "---|";
  std::cout << "\n" << std::setprecision(3);
  for (double pert = 4.0; pert >= min_pert; pert *= 0.5) {
    std::vector<double> err(NV, 0.0);
    for (int r = 0; r < reps; r++) {
      // W = 1 + pert * U[0,1) from the counter RNG, plus the two spikes
      SKINNY_CHECK(skinny_fill_uniform(W.p, (int64_t)n, 77, (uint64_t)r, 0, nullptr));
      HIP_CHECK(hipMemcpy(hW.data(), W.p, sizeof(double) * n, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; i++) hW[i] = 1.0 + pert * 0.5 * (hW[i] + 1.0);
      hW[234] = 100000.0;
      hW[235] = -100000.0;
      for (size_t i = 0; i < n; i++) hWf[i] = (float)hW[i];
      HIP_CHECK(hipMemcpy(W.p, hW.data(), sizeof(double) * n, hipMemcpyHostToDevice));
      HIP_CHECK(hipMemcpy(Wf.p, hWf.data(), sizeof(float) * n, hipMemcpyHostToDevice));
      for (int v = 0; v < DEVICE; v++) {
        gram(v);
        const std::vector<double> hT = svqb_transform(M, hS);
        HIP_CHECK(hipMemcpy(T.p, hT.data(), sizeof(double) * M * M, hipMemcpyHostToDevice));
        SKINNY_CHECK(skinny_dtsmm(nullptr, 0, M, M, K, W.p, M, 1.0, T.p, M, 0.0, Q.p, M,
                                  nullptr));
        err[v] += max_error(Q.p) / reps;
      }
      // the library's device SVQB, one sweep, in place on a copy
      HIP_CHECK(hipMemcpy(Q.p, W.p, sizeof(double) * n, hipMemcpyDeviceToDevice));
      skinny_svqb_block blk{M, K, Q.p, M};
      SKINNY_CHECK(skinny_dsvqb(nullptr, 1, &blk, 1, nullptr));
      err[DEVICE] += max_error(Q.p) / reps;
    }
    std::cout << "| " << pert << " |";
    for (int v = 0; v < NV; v++) std::cout << " " << err[v] << " |";
    std::cout << "\n" << std::flush;
  }
  return 0;
}
