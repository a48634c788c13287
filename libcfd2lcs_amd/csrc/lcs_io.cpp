// lcs_io.cpp -- Fortran-compatible output files of the FTLE path (Tecplot
// ASCII FTLE files and unformatted slot files), written by a background
// thread from host copies of the device results.
#include "lcs_io.h"

#include <sys/stat.h>
#include <sys/types.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace lcs {
namespace io {

// ---------------------------------------------------------------------------
// ES edit descriptor
// ---------------------------------------------------------------------------
static void justify(const char* s, int len, int w, char* out) {
  if (len > w) {
    std::memset(out, '*', (size_t)w);
    return;
  }
  std::memset(out, ' ', (size_t)(w - len));
  std::memcpy(out + (w - len), s, (size_t)len);
}

static int exp_field(int ex, char* p) {
  // E+zz for |e| <= 99, +zzz for |e| <= 999 (Fortran drops the letter)
  const int a = ex < 0 ? -ex : ex;
  const char sg = ex < 0 ? '-' : '+';
  if (a <= 99) {
    p[0] = 'E';
    p[1] = sg;
    p[2] = (char)('0' + a / 10);
    p[3] = (char)('0' + a % 10);
    return 4;
  }
  if (a <= 999) {
    p[0] = sg;
    p[1] = (char)('0' + a / 100);
    p[2] = (char)('0' + (a / 10) % 10);
    p[3] = (char)('0' + a % 10);
    return 4;
  }
  return -1;
}

void format_es(double v, int w, int d, char* out) {
  char buf[96];
  int len = 0;
  if (std::isnan(v)) {
    std::memcpy(buf, "NaN", 3);
    len = 3;
  } else if (std::isinf(v)) {
    const bool neg = std::signbit(v);
    const char* s = neg ? (w >= 9 ? "-Infinity" : "-Inf") : (w >= 8 ? "Infinity" : "Inf");
    len = (int)std::strlen(s);
    std::memcpy(buf, s, (size_t)len);
  } else {
    char t[96];
    std::snprintf(t, sizeof t, "%.*E", d, v);
    char* e = std::strchr(t, 'E');
    const int ex = std::atoi(e + 1);
    len = (int)(e - t);
    std::memcpy(buf, t, (size_t)len);
    const int k = exp_field(ex, buf + len);
    if (k < 0) {
      std::memset(out, '*', (size_t)w);
      return;
    }
    len += k;
  }
  justify(buf, len, w, out);
}

// 10^k as the nearest double, k in [kPmin, kPmax]
namespace {
constexpr int kPmin = -307, kPmax = 308;
struct Pow10 {
  double p[kPmax - kPmin + 1];
  Pow10() {
    char b[16];
    for (int k = kPmin; k <= kPmax; k++) {
      std::snprintf(b, sizeof b, "1e%d", k);
      p[k - kPmin] = std::strtod(b, nullptr);  // correctly rounded
    }
  }
};
const Pow10& pow10_table() {
  static const Pow10 t;
  return t;
}
}  // namespace

void format_es14_6(double v, char* out) {
  if (!std::isfinite(v)) {
    format_es(v, 14, 6, out);
    return;
  }
  const bool neg = std::signbit(v);
  const double a = std::fabs(v);
  long long m = 0;
  int e10 = 0;
  if (a != 0) {
    const Pow10& P = pow10_table();
    // decimal exponent from the binary one: floor(e2 * log10(2)) or one less
    const int e2 = std::ilogb(a);
    e10 = (e2 * 78913) >> 18;
    double s = 0;
    // 7 significant digits: s = a * 10^(6 - e10) in [999999.5, 9999999.5)
    // rounds to m in [1000000, 9999999]; the two bounds overlap in outcome,
    // so a correction never oscillates
    for (int it = 0; it < 4; it++) {
      const int k = 6 - e10;
      if (k < kPmin || k > kPmax) {
        format_es(v, 14, 6, out);
        return;
      }
      s = a * P.p[k - kPmin];
      // a rounding tie of the 7th digit also sits at each decade bound
      if (std::fabs(s - 9999999.5) < 1e-7 || std::fabs(s - 999999.5) < 1e-7) {
        format_es(v, 14, 6, out);
        return;
      }
      if (s < 999999.5) e10--;
      else if (s >= 9999999.5) e10++;
      else break;
    }
    if (!(s >= 999999.5 && s < 9999999.5)) {
      format_es(v, 14, 6, out);
      return;
    }
    // s carries a relative error below 2.3e-16 (the rounded power and the
    // product), i.e. below 3e-9 absolute: near a rounding tie the exact
    // decimal conversion decides
    const double fl = std::floor(s);
    const double fr = s - fl;
    if (std::fabs(fr - 0.5) < 1e-7) {
      format_es(v, 14, 6, out);
      return;
    }
    m = (long long)fl + (fr > 0.5 ? 1 : 0);
    if (m >= 10000000) {
      m /= 10;
      e10++;
    }
  }
  char buf[16];
  int len = 0;
  if (neg) buf[len++] = '-';
  buf[len++] = (char)('0' + (int)(m / 1000000));
  buf[len++] = '.';
  long long r = m % 1000000;
  for (int q = 5; q >= 0; q--) {
    buf[len + q] = (char)('0' + (int)(r % 10));
    r /= 10;
  }
  len += 6;
  const int k = exp_field(e10, buf + len);
  if (k < 0) {
    format_es(v, 14, 6, out);
    return;
  }
  len += k;
  justify(buf, len, 14, out);
}

bool findex_string(long long f, std::string& out) {
  if (f < -999999999LL || f > 999999999LL) return false;
  char b[32];
  if (f >= 1000 && f <= 9999) std::snprintf(b, sizeof b, "%010lld", f);
  else std::snprintf(b, sizeof b, "%lld", f);
  out = b;
  return true;
}

std::string i4(long long v) {
  if (v < 0 || v > 9999) return std::string("****");
  char b[8];
  std::snprintf(b, sizeof b, "%04lld", v);
  return b;
}

// ---------------------------------------------------------------------------
// Tecplot blocks
// ---------------------------------------------------------------------------
size_t format_lines(const Array& a, size_t first, size_t count, char* out) {
  char* p = out;
  const size_t end = first + count;
  for (size_t i = first; i < end; i++) {
    const double v = a.is_double ? ((const double*)a.p)[i] : (double)((const float*)a.p)[i];
    format_es14_6(v, p);
    p += 14;
    if ((i + 1) % 10 == 0 || i + 1 == a.n) *p++ = '\n';
  }
  return (size_t)(p - out);
}

static bool finish(FILE* f, const std::string& tmp, const std::string& path, std::string& err) {
  const bool ok = std::fflush(f) == 0 && !std::ferror(f);
  const bool closed = std::fclose(f) == 0;
  if (!ok || !closed) {
    err = "write error on " + tmp + ": " + std::strerror(errno);
    std::remove(tmp.c_str());
    return false;
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) {
    err = "cannot rename " + tmp + " to " + path + ": " + std::strerror(errno);
    return false;
  }
  return true;
}

bool write_tecplot(const Job& j, int nthreads, std::string& err) {
  const std::string tmp = j.path + ".part";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) {
    err = "cannot open " + tmp + ": " + std::strerror(errno);
    return false;
  }
  std::fwrite(j.header.data(), 1, j.header.size(), f);
  nthreads = std::max(1, nthreads);
  const size_t per = 100000;  // values per thread and slab, a multiple of 10
  std::vector<std::vector<char>> bufs((size_t)nthreads, std::vector<char>(per * 14 + per / 10 + 16));
  std::vector<size_t> used((size_t)nthreads);
  for (const Array& a : j.arrays) {
    for (size_t s0 = 0; s0 < a.n; s0 += per * (size_t)nthreads) {
      const size_t s1 = std::min(a.n, s0 + per * (size_t)nthreads);
      const int parts = (int)((s1 - s0 + per - 1) / per);
      auto work = [&](int t) {
        const size_t b = s0 + (size_t)t * per;
        const size_t e = std::min(s1, b + per);
        used[(size_t)t] = format_lines(a, b, e - b, bufs[(size_t)t].data());
      };
      if (parts == 1) {
        work(0);
      } else {
        std::vector<std::thread> ts;
        for (int t = 1; t < parts; t++) ts.emplace_back(work, t);
        work(0);
        for (auto& t : ts) t.join();
      }
      for (int t = 0; t < parts; t++) std::fwrite(bufs[(size_t)t].data(), 1, used[(size_t)t], f);
    }
  }
  return finish(f, tmp, j.path, err);
}

// gfortran's unformatted sequential records: a 4-byte length before and
// after the data; records longer than 2147483639 bytes are split into
// subrecords whose leading marker is negative when another subrecord follows
// and whose trailing marker is negative when one precedes.
bool write_unformatted(const Job& j, std::string& err) {
  const std::string tmp = j.path + ".part";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) {
    err = "cannot open " + tmp + ": " + std::strerror(errno);
    return false;
  }
  const size_t kMaxSub = 2147483639u;
  for (const Array& a : j.arrays) {
    const size_t bytes = a.n * (a.is_double ? 8 : 4);
    const char* p = (const char*)a.p;
    size_t left = bytes;
    bool firstsub = true;
    do {
      const size_t chunk = std::min(left, kMaxSub);
      const bool more = left > chunk;
      const int32_t head = more ? -(int32_t)chunk : (int32_t)chunk;
      const int32_t tail = firstsub ? (int32_t)chunk : -(int32_t)chunk;
      std::fwrite(&head, 4, 1, f);
      if (chunk) std::fwrite(p, 1, chunk, f);
      std::fwrite(&tail, 4, 1, f);
      p += chunk;
      left -= chunk;
      firstsub = false;
    } while (left > 0);
  }
  return finish(f, tmp, j.path, err);
}

bool ensure_dir(const std::string& dir, std::string& err) {
  struct stat sb;
  if (stat(dir.c_str(), &sb) == 0) {
    if (S_ISDIR(sb.st_mode)) return true;
    err = dir + " exists and is not a directory";
    return false;
  }
  if (mkdir(dir.c_str(), 0777) != 0 && errno != EEXIST) {
    err = "cannot create the output directory " + dir + ": " + std::strerror(errno);
    return false;
  }
  return true;
}

// ---------------------------------------------------------------------------
// background writer
// ---------------------------------------------------------------------------
Writer::Writer(int max_pending) : max_pending_(std::max(1, max_pending)) {
  const unsigned hc = std::thread::hardware_concurrency();
  nthreads_ = (int)std::min(16u, std::max(1u, hc));
  th_ = std::thread([this] { run(); });
}

Writer::~Writer() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (th_.joinable()) th_.join();
}

void Writer::submit(Job&& j) {
  std::unique_lock<std::mutex> lk(mu_);
  cv_.wait(lk, [&] { return (int)q_.size() < max_pending_; });
  q_.push_back(std::move(j));
  lk.unlock();
  cv_.notify_all();
}

void Writer::flush() {
  std::unique_lock<std::mutex> lk(mu_);
  cv_.wait(lk, [&] { return q_.empty() && busy_ == 0; });
}

bool Writer::failed(std::string* msg) {
  std::lock_guard<std::mutex> lk(mu_);
  if (err_ && msg) *msg = msg_;
  return err_;
}

void Writer::run() {
  for (;;) {
    Job j;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || !q_.empty(); });
      if (q_.empty()) return;  // stop requested and drained
      j = std::move(q_.front());
      q_.pop_front();
      busy_++;
    }
    cv_.notify_all();
    std::string e;
    const bool ok = j.kind == Job::kTecplot ? write_tecplot(j, nthreads_, e) : write_unformatted(j, e);
    j = Job();  // release the host copies before reporting completion
    {
      std::lock_guard<std::mutex> lk(mu_);
      busy_--;
      if (!ok && !err_) {
        err_ = true;
        msg_ = e;
      }
    }
    cv_.notify_all();
  }
}

}  // namespace io
}  // namespace lcs
