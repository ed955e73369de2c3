// Host Hermitian/symmetric eigensolver (rcg_dense.hpp small_eigensystem) on a
// matrix read from a binary file: n (int64), then n*n complex<double>
// (column-major).  Writes n eigenvalues and the n*n complex eigenvectors
// (column-major) to the output file.  Used by tests/test_dense_eig_host.py.
#include <complex>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../rayleighchebyshevgpgpu_amd/csrc/rcg_dense.hpp"

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  FILE* in = std::fopen(argv[1], "rb");
  if (!in) return 3;
  int64_t n = 0;
  if (std::fread(&n, sizeof(n), 1, in) != 1 || n <= 0) return 4;
  std::vector<std::complex<double>> H(n * n);
  if (std::fread(H.data(), sizeof(std::complex<double>), H.size(), in) != H.size()) return 5;
  std::fclose(in);
  std::vector<double> w;
  std::vector<std::complex<double>> Z;
  rcg::dense::small_eigensystem(H.data(), (long)n, w, Z);
  FILE* out = std::fopen(argv[2], "wb");
  if (!out) return 6;
  std::fwrite(w.data(), sizeof(double), w.size(), out);
  std::fwrite(Z.data(), sizeof(std::complex<double>), Z.size(), out);
  std::fclose(out);
  return 0;
}
