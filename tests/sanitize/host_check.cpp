// Host-side checks run under AddressSanitizer and UndefinedBehaviorSanitizer
// (built host-only by tests/test_host_sanitizers.py). Commands on stdin, one
// per line, results on stdout:
//   field                 self-check of the host field arithmetic (BabyBear,
//                         BabyBear^4, Goldilocks) against 128-bit integer
//                         references, edge operands included: prints "ok"
//   sha <hex>             SHA-256 digest of the bytes (the library's stream)
//   b3 <hex> <split> <n>  BLAKE3 XOF, n bytes, of the bytes absorbed in two
//                         parts split at <split>
//   hop <kind> <item> <limbs> <hex-prefix> <hex-msg>
//                         one Fiat-Shamir hop of the field transcript after
//                         absorbing the prefix: prints the squeezed bytes
//   oref <path>           a zerocheck proof of the C oracle on the request in
//                         <path>; prints the messages, point and value
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../zorch_amd/csrc/blake3.h"
#include "../../zorch_amd/csrc/sha256.h"

extern "C" {
typedef struct {
  uint32_t h[8];
  uint8_t pending[64];
  uint64_t pending_len;
  uint64_t total_len;
} oref_sha_state;
void oref_sha_init(oref_sha_state* s);
void oref_sha_absorb(oref_sha_state* s, const uint8_t* d, size_t n);
int oref_prove(int tab_field, int chal_field, int tr_field, int eq, int degree, int k, int n,
               const void* tables, const void* eq_point, const void* claim_in,
               oref_sha_state* tr, void* msgs_out, void* point_out, void* value_out,
               void* final_out, int threads);
int oref_claim(int tab_field, int chal_field, int eq, int k, int n, const void* tables,
               const void* eq_point, void* claim_out, int threads);
}

using namespace zs;

static std::vector<uint8_t> unhex(const std::string& h) {
  std::vector<uint8_t> out;
  if (h == "-") return out;
  for (size_t i = 0; i + 1 < h.size(); i += 2) out.push_back((uint8_t)strtoul(h.substr(i, 2).c_str(), nullptr, 16));
  return out;
}

static void puthex(const uint8_t* p, size_t n) {
  for (size_t i = 0; i < n; ++i) printf("%02x", p[i]);
  printf("\n");
}

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;
static uint64_t next64() {
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return rng_state;
}

static const uint64_t P_BB = (1ull << 31) - (1ull << 27) + 1;
static const unsigned __int128 P_GL = ((unsigned __int128)1 << 64) - ((unsigned __int128)1 << 32) + 1;

static uint64_t pow_mod(uint64_t b, uint64_t e) {
  uint64_t r = 1;
  b %= P_BB;
  while (e) { if (e & 1) r = r * b % P_BB; b = b * b % P_BB; e >>= 1; }
  return r;
}
static const uint64_t RINV_BB = pow_mod((1ull << 32) % P_BB, P_BB - 2);
static uint64_t bb_val(uint32_t w) { return (uint64_t)w % P_BB * RINV_BB % P_BB; }  // Montgomery word -> value

static int check_fields() {
  int bad = 0;
  for (int it = 0; it < 20000; ++it) {
    // BabyBear: Montgomery words, any u32 in [0, p).
    uint32_t a = (uint32_t)(next64() % P_BB), b = (uint32_t)(next64() % P_BB);
    if (it < 4) { a = (uint32_t)(P_BB - 1 - it); b = (uint32_t)(P_BB - 1); }
    const uint64_t va = bb_val(a), vb = bb_val(b);
    if (bb_val(BB::mul(a, b)) != va * vb % P_BB) ++bad;
    if (bb_val(BB::add(a, b)) != (va + vb) % P_BB) ++bad;
    if (bb_val(BB::sub(a, b)) != (va + P_BB - vb) % P_BB) ++bad;
    if (a % P_BB && bb_val(BB::mul(a, BB::inv(a))) != 1) ++bad;
    // Goldilocks: canonical words; operands near p and 2^64 reach the
    // reduction's borrow, carry and final subtraction.
    uint64_t x = next64() % (uint64_t)P_GL, y = next64() % (uint64_t)P_GL;
    if (it % 7 == 0) x = (uint64_t)P_GL - 1 - (next64() & 0xffff);
    if (it % 11 == 0) y = 0xffffffffull + (next64() & 0xff);
    const unsigned __int128 ref = (unsigned __int128)x * y % P_GL;
    if ((unsigned __int128)GL::canon(GL::mul(x, y)) != ref) ++bad;
    if ((unsigned __int128)GL::canon(GL::add(x, y)) != ((unsigned __int128)x + y) % P_GL) ++bad;
    if (x && GL::canon(GL::mul(x, GL::inv(x))) != 1) ++bad;
    // BabyBear^4 = F_p[X]/(X^4 - 11): product by schoolbook with 128-bit sums.
    bb4 u = bb4_make((uint32_t)(next64() % P_BB), (uint32_t)(next64() % P_BB),
                     (uint32_t)(next64() % P_BB), (uint32_t)(next64() % P_BB));
    bb4 v = bb4_make((uint32_t)(next64() % P_BB), (uint32_t)(next64() % P_BB),
                     (uint32_t)(next64() % P_BB), (uint32_t)(next64() % P_BB));
    uint64_t uu[4], vv[4], prod[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) { uu[i] = bb_val(u.c[i]); vv[i] = bb_val(v.c[i]); }
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) prod[i + j] = (prod[i + j] + uu[i] * vv[j]) % P_BB;
    bb4 w = BB4::mul(u, v);
    for (int i = 0; i < 4; ++i) {
      const uint64_t want = (prod[i] + (i + 4 < 7 ? 11 * prod[i + 4] : 0)) % P_BB;
      if (bb_val(w.c[i]) != want) ++bad;
    }
    bb4 winv = BB4::mul(u, BB4::inv(u));
    if (bb_val(winv.c[0]) != 1 || bb_val(winv.c[1]) || bb_val(winv.c[2]) || bb_val(winv.c[3])) ++bad;
  }
  return bad;
}

int main() {
  char line[1 << 16];
  while (fgets(line, sizeof line, stdin)) {
    std::vector<std::string> tok;
    for (char* t = strtok(line, " \n"); t; t = strtok(nullptr, " \n")) tok.push_back(t);
    if (tok.empty()) continue;
    if (tok[0] == "field") {
      const int bad = check_fields();
      printf(bad ? "bad %d\n" : "ok\n", bad);
    } else if (tok[0] == "sha") {
      std::vector<uint8_t> m = unhex(tok[1]);
      Sha256State s;
      sha256_init(s);
      sha256_absorb(s, m.data(), m.size());
      uint8_t out[32];
      sha256_digest_suffix(s, nullptr, 0, out);
      puthex(out, 32);
    } else if (tok[0] == "b3") {
      std::vector<uint8_t> m = unhex(tok[1]);
      const size_t split = strtoul(tok[2].c_str(), nullptr, 10), n = strtoul(tok[3].c_str(), nullptr, 10);
      Blake3State s;
      b3_init(s);
      bool ok = b3_absorb(s, m.data(), split) && b3_absorb(s, m.data() + split, m.size() - split);
      std::vector<uint8_t> out(n);
      b3_finalize(s, out.data(), n);
      if (!ok) printf("overflow\n");
      else puthex(out.data(), n);
    } else if (tok[0] == "hop") {
      const int kind = atoi(tok[1].c_str());
      const size_t item = strtoul(tok[2].c_str(), nullptr, 10), limbs = strtoul(tok[3].c_str(), nullptr, 10);
      std::vector<uint8_t> pre = unhex(tok[4]), msg = unhex(tok[5]);
      std::vector<uint8_t> out(item * limbs);
      if (kind == 0) {
        Sha256State s;
        sha256_init(s);
        sha256_absorb(s, pre.data(), pre.size());
        sha256_fs_hop(s, msg.data(), msg.size(), item, limbs, out.data());
      } else {
        Blake3State s;
        b3_init(s);
        b3_absorb(s, pre.data(), pre.size());
        b3_fs_hop(s, msg.data(), msg.size(), item, limbs, out.data());
      }
      puthex(out.data(), out.size());
    } else if (tok[0] == "oref") {
      // Request: n (u32), two BabyBear tables of 2^n words, n BabyBear^4
      // eq-point words (4 a coordinate), the transcript stream bytes.
      FILE* f = fopen(tok[1].c_str(), "rb");
      if (!f) { printf("no file\n"); continue; }
      uint32_t n = 0;
      if (fread(&n, 4, 1, f) != 1 || n < 1 || n > 16) { printf("bad n\n"); fclose(f); continue; }
      const size_t N = (size_t)1 << n;
      std::vector<uint32_t> tabs(2 * N), z(4 * n);
      size_t got = fread(tabs.data(), 4, tabs.size(), f);
      got += fread(z.data(), 4, z.size(), f);
      std::vector<uint8_t> stream(1 << 12);
      const size_t slen = fread(stream.data(), 1, stream.size(), f);
      fclose(f);
      if (got != tabs.size() + z.size()) { printf("short file\n"); continue; }
      oref_sha_state st;
      oref_sha_init(&st);
      oref_sha_absorb(&st, stream.data(), slen);
      std::vector<uint32_t> claim(4), msgs((size_t)n * 4 * 4), point((size_t)n * 4), value(4),
          fin(3 * 4);
      // ids: 1 BabyBear, 2 BabyBear^4 (oracle/sumcheck_ref.c)
      if (oref_claim(1, 2, 1, 2, (int)n, tabs.data(), z.data(), claim.data(), 1) ||
          oref_prove(1, 2, 1, 1, 2, 2, (int)n, tabs.data(), z.data(), claim.data(), &st,
                     msgs.data(), point.data(), value.data(), fin.data(), 1)) {
        printf("prove failed\n");
        continue;
      }
      puthex((const uint8_t*)msgs.data(), msgs.size() * 4);
      puthex((const uint8_t*)point.data(), point.size() * 4);
      puthex((const uint8_t*)value.data(), value.size() * 4);
    }
    fflush(stdout);
  }
  return 0;
}
