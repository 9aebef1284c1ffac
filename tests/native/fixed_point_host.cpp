// Host harness of the wide fixed-point helpers in device_util.h (built by
// tests/test_fixed_point.py with hipcc, host code only). Reads one request
// per line on stdin and answers each on stdout:
//   fix <bits hex> <S> <N>          the N-word image trunc(x * 2^S)
//   sq <bits hex> <S2> <N>          the N-word image trunc(x^2 * 2^S2)
//   tof <n> <scale> <words hex...>  an n-word two's-complement value * 2^-scale
//   var <nf> <S> <nq> <S2> <n> <ddof> <sum words> <sq words>
//   var2 <S> <S2> <n> <ddof> <sum lo hi> <sq lo hi>   the two-word finish
//   u256 <scale> <4 words hex>      a non-negative 4-word value * 2^-scale
//   shl <s> <4 words hex>           the 4-word value shifted left by s
#include "../../dataframe_mojo_amd/csrc/device_util.h"
#include <cstdio>
#include <cstdlib>
#include <cstring>
using namespace dfr;
// reads lines: op args ; prints results
int main() {
  char op[16];
  while (scanf("%15s", op) == 1) {
    if (!strcmp(op, "fix")) {  // fix bits S N -> words
      unsigned long long bits; int S, N;
      scanf("%llx %d %d", &bits, &S, &N);
      uint64_t w[4];
      fix_words<4>(bits, S, w);
      for (int i = 0; i < N; ++i) printf("%016llx ", (unsigned long long)w[i]);
      printf("\n");
    } else if (!strcmp(op, "sq")) {
      unsigned long long bits; int S2, N;
      scanf("%llx %d %d", &bits, &S2, &N);
      uint64_t w[4];
      fixsq_words<4>(bits, S2, w);
      for (int i = 0; i < N; ++i) printf("%016llx ", (unsigned long long)w[i]);
      printf("\n");
    } else if (!strcmp(op, "tof")) {  // tof n scale words...
      int n, scale; scanf("%d %d", &n, &scale);
      uint64_t w[4] = {0,0,0,0};
      for (int i = 0; i < n; ++i) { unsigned long long x; scanf("%llx", &x); w[i] = x; }
      double d = sword_to_f64(w, n, scale);
      uint64_t b; memcpy(&b, &d, 8);
      printf("%016llx\n", (unsigned long long)b);
    } else if (!strcmp(op, "var")) {  // var nf S nq S2 n ddof sumwords sqwords
      int nf, S, nq, S2; unsigned long long n, dd;
      scanf("%d %d %d %d %llu %llu", &nf, &S, &nq, &S2, &n, &dd);
      uint64_t a[4] = {0}, b[4] = {0};
      for (int i = 0; i < nf; ++i) { unsigned long long x; scanf("%llx", &x); a[i] = x; }
      for (int i = 0; i < nq; ++i) { unsigned long long x; scanf("%llx", &x); b[i] = x; }
      double d = var_words(a, nf, S, b, nq, S2, n, dd);
      uint64_t bb; memcpy(&bb, &d, 8);
      printf("%016llx\n", (unsigned long long)bb);
    } else if (!strcmp(op, "var2")) {
      int S, S2; unsigned long long n, dd, tl, th, ql, qh;
      scanf("%d %d %llu %llu %llx %llx %llx %llx", &S, &S2, &n, &dd, &tl, &th, &ql, &qh);
      double d = var_2words(tl, th, S, ql, qh, S2, n, dd);
      uint64_t bb; memcpy(&bb, &d, 8);
      printf("%016llx\n", (unsigned long long)bb);
    } else if (!strcmp(op, "u256") || !strcmp(op, "shl")) {
      int k; scanf("%d", &k);
      U256 x;
      for (int i = 0; i < 4; ++i) { unsigned long long v; scanf("%llx", &v); x.w[i] = v; }
      if (op[0] == 'u') {
        double d = u256_to_f64(x, k);
        uint64_t bb; memcpy(&bb, &d, 8);
        printf("%016llx\n", (unsigned long long)bb);
      } else {
        U256 r = shl256(x, k);
        for (int i = 0; i < 4; ++i) printf("%016llx ", (unsigned long long)r.w[i]);
        printf("\n");
      }
    }
    fflush(stdout);
  }
}
