// Density restart files (MOPAC's .den): the alpha density of a closed-shell
// SCF as den_in_out writes it (src/SCF/iter.F90:1133-1227) and OLDENS reads
// it (iter.F90:151-159), so the library's SCF and MOPAC can start from each
// other's densities.
//
// Layout, as the Fortran statements `write (iden) norbs, numat, pa` and
// `write (iden, *) norbs, numat, pa` produce it:
//   unformatted  one sequential record: a 4-byte length marker, norbs and
//                numat (4-byte integers), pa (mpack 8-byte reals), the same
//                marker again; records over 2^31 - 1 bytes are split into
//                subrecords whose leading markers are negated except on the
//                last, and whose trailing markers are negated except on the
//                first (the Fortran runtime's convention);
//   formatted    list-directed text: norbs, numat, then the mpack values.
// Reading tries the form asked for first and then the other, as den_in_out
// does, and rejects a file whose norbs or numat differ from the system's.
#include <cerrno>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mopac_hip.h"

namespace mhip {
void set_error(const std::string& msg);  // the message mhip_last_error() returns
}

namespace {

// The message mhip_last_error() returns, kept here too so a read can fall
// back to the first form's message.
thread_local std::string g_den_err;

int fail(const std::string& msg) {
  g_den_err = msg;
  mhip::set_error(msg);
  return 1;
}

constexpr int64_t kMaxSub = 2147483639;  // the runtime's largest subrecord payload

bool write_unformatted(FILE* f, const std::vector<char>& payload) {
  const int64_t total = (int64_t)payload.size();
  int64_t done = 0;
  bool first = true;
  do {
    const int64_t len = std::min<int64_t>(kMaxSub, total - done);
    const bool last = (done + len == total);
    const int32_t lead = (int32_t)(last ? len : -len);
    const int32_t tail = (int32_t)(first ? len : -len);
    if (std::fwrite(&lead, 4, 1, f) != 1) return false;
    if (len && std::fwrite(payload.data() + done, 1, (size_t)len, f) != (size_t)len) return false;
    if (std::fwrite(&tail, 4, 1, f) != 1) return false;
    done += len;
    first = false;
  } while (done < total);
  return true;
}

bool read_unformatted(FILE* f, std::vector<char>& payload) {
  payload.clear();
  for (;;) {
    int32_t lead = 0, tail = 0;
    if (std::fread(&lead, 4, 1, f) != 1) return false;
    const bool more = lead < 0;
    const int64_t len = more ? -(int64_t)lead : lead;
    const size_t at = payload.size();
    payload.resize(at + (size_t)len);
    if (len && std::fread(payload.data() + at, 1, (size_t)len, f) != (size_t)len) return false;
    if (std::fread(&tail, 4, 1, f) != 1) return false;
    if ((tail < 0 ? -(int64_t)tail : tail) != len) return false;
    if (!more) return true;
  }
}

int read_file(const char* path, bool formatted, int norbs, int numat, double* pa, long long mpack) {
  FILE* f = std::fopen(path, formatted ? "r" : "rb");
  if (!f) return fail(std::string("Density Restart File missing or corrupt: cannot open ") + path);
  int old_norbs = 0, old_numat = 0;
  bool ok = false;
  if (formatted) {
    ok = std::fscanf(f, "%d %d", &old_norbs, &old_numat) == 2;
    for (long long q = 0; ok && q < mpack;) {
      // list-directed values: "v" or "r*v" (r repeats), a D exponent allowed
      char tok[96];
      ok = std::fscanf(f, " %95[^ \t\r\n,]", tok) == 1;
      if (!ok) break;
      long long rep = 1;
      char* val = tok;
      if (char* star = std::strchr(tok, '*')) {
        *star = '\0';
        rep = std::strtoll(tok, nullptr, 10);
        val = star + 1;
      }
      for (char* c = val; *c; ++c)
        if (*c == 'D' || *c == 'd') *c = 'E';
      char* end = nullptr;
      errno = 0;
      const double v = std::strtod(val, &end);
      ok = end != val && errno == 0 && rep >= 1 && q + rep <= mpack;
      for (long long r = 0; ok && r < rep; ++r) pa[q++] = v;
      int ch = std::fgetc(f);  // one separating comma, if any
      if (ch != ',' && ch != EOF) std::ungetc(ch, f);
    }
  } else {
    // the header is taken even from a record too short for pa, so that a
    // size mismatch is reported as such (den_in_out tests the header first)
    std::vector<char> rec;
    ok = read_unformatted(f, rec);
    if (rec.size() >= 8) {
      int32_t hdr[2];
      std::memcpy(hdr, rec.data(), 8);
      old_norbs = hdr[0];
      old_numat = hdr[1];
    }
    ok = ok && (int64_t)rec.size() >= 8 + 8 * mpack;
    if (ok) std::memcpy(pa, rec.data() + 8, sizeof(double) * (size_t)mpack);
  }
  std::fclose(f);
  if (old_norbs > 0 && old_norbs < 100000 && (old_norbs != norbs || old_numat != numat))
    return fail("Density file read in does not match current data set");
  if (!ok) return fail("Density Restart File missing or corrupt");
  return 0;
}

}  // namespace

extern "C" {

int mhip_den_write(const char* path, int norbs, int numat, const double* pa, int formatted) {
  if (!path || norbs <= 0 || numat <= 0 || !pa) return fail("mhip_den_write: bad arguments");
  const long long mpack = (long long)norbs * (norbs + 1) / 2;
  FILE* f = std::fopen(path, formatted ? "w" : "wb");
  if (!f) return fail(std::string("mhip_den_write: cannot open ") + path);
  bool ok;
  if (formatted) {
    ok = std::fprintf(f, " %11d %11d\n", norbs, numat) > 0;
    for (long long q = 0; ok && q < mpack; ++q)
      ok = std::fprintf(f, " %24.16E%s", pa[q], (q % 3 == 2 || q == mpack - 1) ? "\n" : "") > 0;
  } else {
    std::vector<char> rec(8 + 8 * (size_t)mpack);
    const int32_t hdr[2] = {norbs, numat};
    std::memcpy(rec.data(), hdr, 8);
    std::memcpy(rec.data() + 8, pa, sizeof(double) * (size_t)mpack);
    ok = write_unformatted(f, rec);
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : fail(std::string("mhip_den_write: write failed on ") + path);
}

int mhip_den_read(const char* path, int norbs, int numat, double* pa, int formatted) {
  if (!path || norbs <= 0 || numat <= 0 || !pa) return fail("mhip_den_read: bad arguments");
  const long long mpack = (long long)norbs * (norbs + 1) / 2;
  // the form asked for, then the other, as den_in_out tries them
  if (read_file(path, formatted != 0, norbs, numat, pa, mpack) == 0) return 0;
  const std::string first = g_den_err;
  if (first.find("does not match") != std::string::npos) return 1;
  if (read_file(path, formatted == 0, norbs, numat, pa, mpack) == 0) return 0;
  if (g_den_err.find("does not match") == std::string::npos) fail(first);
  return 1;
}

}  // extern "C"
