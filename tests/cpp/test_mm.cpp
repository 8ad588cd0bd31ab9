// MatrixMarket reader driver (CPU only): prints what the facade read, for
// the Python test to compare with scipy.io.mmread.
//   test_mm matrix <file>   -> "n nnz", then one "i j value" line per CSR entry
//   test_mm vector <file>   -> "n", then one value per line
//   test_mm roundtrip <in> <out>  -> reads a vector, writes it, reads it back
// Errors: "ERROR <message>" on stdout and exit status 2.
#include <cstdio>
#include <string>

#include "MCLS_MatrixMarket.hpp"

using namespace MCLS;

int main(int argc, char** argv) {
  if (argc < 3) {
    std::printf("usage: test_mm matrix|vector|roundtrip <file> [out]\n");
    return 1;
  }
  const std::string mode = argv[1];
  try {
    if (mode == "matrix") {
      MatrixMarketInfo info;
      auto A = readMatrixMarket<double>(argv[2], &info);
      std::printf("%lld %lld\n", (long long)A->numRows(), (long long)A->nnz());
      for (int64_t i = 0; i < A->numRows(); ++i)
        for (int64_t k = A->rowPtr()[(size_t)i]; k < A->rowPtr()[(size_t)i + 1]; ++k)
          std::printf("%lld %d %.17g\n", (long long)i, A->colInd()[(size_t)k], A->values()[(size_t)k]);
    } else if (mode == "vector") {
      auto v = readMatrixMarketVector(argv[2]);
      std::printf("%lld\n", (long long)v->size());
      for (int64_t i = 0; i < v->size(); ++i) std::printf("%.17g\n", (*v)[i]);
    } else if (mode == "roundtrip" && argc == 4) {
      auto v = readMatrixMarketVector(argv[2]);
      writeMatrixMarketVector(argv[3], *v);
      auto w = readMatrixMarketVector(argv[3]);
      bool same = v->size() == w->size();
      for (int64_t i = 0; same && i < v->size(); ++i) same = (*v)[i] == (*w)[i];
      std::printf(same ? "same\n" : "different\n");
      return same ? 0 : 3;
    } else {
      std::printf("bad mode\n");
      return 1;
    }
  } catch (const Assertion& e) {
    std::printf("ERROR %s\n", e.what());
    return 2;
  }
  return 0;
}
