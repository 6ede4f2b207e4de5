// Drop-in check of the reference's host driver surface: the call lines below are the
// reference's own, compiled unchanged against include/ht_amd.h and linked with
// libht_amd.so. The rest of each reference driver (cudaHostAlloc, NVML queries, the
// cuBLAS/cuSOLVER handle creation, genHtensor's unseeded input) is replaced by reading a
// seeded tensor from a file, so that the printed errors can be checked.
//
//   driver <variant> <tensor.f32> <n1> <n2> <n3> [<n4>] <k0> <k1> ... <k_{2d-2}>
//   variant: 3B | 3O | 3T | 4B | 4O
//
// 3B and 3O print "driver error <value>" with the returned error; every variant's library
// call prints the reference's own "error rate ..." line.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/ht_amd.h"

typedef float dt;  // the reference's scalar type (function.h:4 of each driver)

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const char* v = argv[1];
  const bool four = v[0] == '4';
  const int order = four ? 4 : 3;
  const int nk = 2 * order - 1;
  if (argc != 3 + order + nk) return 2;
  long n[4] = {1, 1, 1, 1};
  for (int i = 0; i < order; ++i) n[i] = atol(argv[3 + i]);
  int* k = new int[nk]();
  for (int i = 0; i < nk; ++i) k[i] = atoi(argv[3 + order + i]);
  const long a = n[0], b = n[1], c = n[2], d = n[3];
  const size_t count = (size_t)a * b * c * d;
  std::vector<dt> buf(count);
  FILE* f = fopen(argv[2], "rb");
  if (!f || fread(buf.data(), sizeof(dt), count, f) != count) return 3;
  fclose(f);
  dt* X = buf.data();
  float error = -1.0f;
  dt rel_eps = 0;       // the reference's 3B driver passes 1e-7/2 and its mode size; both are
  dt max_rank = 0;      // ignored there, as here when rel_eps is not positive
  void* handle = nullptr;     // cublasHandle_t in the reference; unused on this platform
  void* cusolverH = nullptr;  // cusolverDnHandle_t in the reference; unused
  if (!strcmp(v, "3B")) {
    error = htd(X,a,b,c,k,rel_eps,max_rank);  // third_order_tensor_decompsition/CUDA/Baseline/ht_init.cu:45
    printf("driver error %.9g\n", error);
  } else if (!strcmp(v, "3O")) {
    error = htd(X,a,b,c,k);                   // third_order_tensor_decompsition/CUDA/Optimized/ht_init.cu:47
    printf("driver error %.9g\n", error);
  } else if (!strcmp(v, "3T")) {
    htd(X,a,b,c,k);                           // third_order_tensor_decompsition/CUDA/TSQR/ht_init.cu:35
  } else if (!strcmp(v, "4B")) {
    htd4(X,a,b,c,d,k);                        // fourth_order_tensor_decomposition/single_GPU/CUDA/Baseline/ht_init.cu:29
  } else if (!strcmp(v, "4O")) {
    htd4(X,a,b,c,d,k,handle,cusolverH);       // fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/ht_init.cu:53
  } else {
    return 2;
  }
  fflush(stdout);
  delete[] k;
  return 0;
}
