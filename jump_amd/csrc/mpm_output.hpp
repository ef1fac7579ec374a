// mpm_output.hpp -- host-side output of the particle state: Julia's text
// form of Float64 and the LAMMPS particle dump of the reference's
// OvitoOutput (Output.jl:146-260), written by a worker thread from pinned
// host copies so that the device keeps stepping while the text is formatted.
//
// Host code only (no device code, no HIP types): the library includes it,
// and tests/host builds it with g++ to check the formatter.
#pragma once

#include <charconv>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace jh {

// ------------------------------------------------------------------ numbers
// print(io, x::Float64) as Julia writes it (its Ryu shortest form): the
// shortest digits that read back as x; plain notation when the decimal point
// falls after at most 6 digits and before at most 4 leading zeros
// (-4 < pt <= 6, pt = decimal exponent + 1), else d.ddd e X with the
// exponent unpadded and unsigned when positive; integers keep ".0".
//   1.0e6   100000.0   0.0001   1.0e-5   1.2345678e6   -0.0   NaN   Inf
// Returns the number of characters written to buf (at least 32 bytes).
inline int julia_float(double x, char* buf) {
  if (std::isnan(x)) { std::memcpy(buf, "NaN", 3); return 3; }
  if (std::isinf(x)) {
    if (x > 0) { std::memcpy(buf, "Inf", 3); return 3; }
    std::memcpy(buf, "-Inf", 4);
    return 4;
  }
  int n = 0;
  if (std::signbit(x)) { buf[n++] = '-'; x = -x; }
  if (x == 0.0) { std::memcpy(buf + n, "0.0", 3); return n + 3; }
  // shortest round-trip digits in scientific form: d[.ddd]e[+-]XX
  char sci[40];
  const auto res = std::to_chars(sci, sci + sizeof(sci), x, std::chars_format::scientific);
  const int len = int(res.ptr - sci);
  int epos = 0;
  while (sci[epos] != 'e') epos++;
  char digits[24];
  int nd = 0;
  for (int i = 0; i < epos; i++)
    if (sci[i] != '.') digits[nd++] = sci[i];
  int e10 = 0;
  std::from_chars(sci + epos + (sci[epos + 1] == '+' ? 2 : 1), sci + len, e10);
  const int pt = e10 + 1;  // digits before the decimal point
  if (pt > -4 && pt <= 6) {
    if (pt <= 0) {  // 0.000ddd
      buf[n++] = '0';
      buf[n++] = '.';
      for (int i = 0; i < -pt; i++) buf[n++] = '0';
      for (int i = 0; i < nd; i++) buf[n++] = digits[i];
    } else if (pt >= nd) {  // ddd000.0
      for (int i = 0; i < nd; i++) buf[n++] = digits[i];
      for (int i = nd; i < pt; i++) buf[n++] = '0';
      buf[n++] = '.';
      buf[n++] = '0';
    } else {  // dd.ddd
      for (int i = 0; i < pt; i++) buf[n++] = digits[i];
      buf[n++] = '.';
      for (int i = pt; i < nd; i++) buf[n++] = digits[i];
    }
    return n;
  }
  buf[n++] = digits[0];
  buf[n++] = '.';
  if (nd == 1) buf[n++] = '0';
  for (int i = 1; i < nd; i++) buf[n++] = digits[i];
  buf[n++] = 'e';
  const auto r2 = std::to_chars(buf + n, buf + n + 8, e10);
  return int(r2.ptr - buf);
}

// ------------------------------------------------------------------ dumps
// Columns after "id type x y z" (Output.jl:178-212 in 2D, 258-281 in 3D).
enum DumpCol : int {
  kColPstrain,   // equivalent plastic strain; "0" for elastic solids
  kColVonMises,  // von Mises stress; "0" for elastic solids
  kColSxx,       // 2D only
  kColSyy,       // 2D only
  kColPressure,  // trace of sigma (2D: s11 + s22)
  kColVx,
  kColVy,
  kColVz,  // 3D only
};

// One queued dump: host copies of the state.  Field f of particle p (the
// pair layout of the particle arrays: 0..D-1 x, D m, D+1 V, D+2..2D+1 v,
// then sigma as 11 22 12 / 11 22 33 12 13 23) is pr[f / 2][2 p + f % 2];
// only the pairs the columns need are set.
struct DumpJob {
  std::string path;
  int dim = 2;
  int64_t np = 0;
  double lims[3] = {0.0, 0.0, 0.0};
  std::vector<std::string> names;  // column names as given
  std::vector<int> cols;
  const double* pr[7] = {};
  const double* alpha = nullptr;
  const double* vm = nullptr;
  std::vector<int64_t> solid_end;  // one past each solid's last particle
  std::vector<uint8_t> elastic;    // per solid: the plastic columns print "0"
  std::function<bool()> ready;     // blocks until the device copy has landed
  std::function<void()> release;   // returns the host buffers
};

// Formats one dump as the reference writes it and appends it to j.path
// (Output.jl opens the file with "a").  Returns false on an I/O error.
inline bool write_dump(const DumpJob& j) {
  std::FILE* f = std::fopen(j.path.c_str(), "ab");
  if (!f) return false;
  std::string out;
  out.reserve(size_t(1) << 21);
  char nb[40];
  auto num = [&](double x) { out.append(nb, size_t(julia_float(x, nb))); };
  const int D = j.dim;
  auto fld = [&](int fi, int64_t p) { return j.pr[fi / 2][2 * p + fi % 2]; };
  out += "ITEM: TIMESTEP \n0\nITEM: NUMBER OF ATOMS\n";
  out += std::to_string(j.np);
  out += " \nITEM: BOX BOUNDS sm sm sm\n";
  for (int a = 0; a < 3; a++) {
    out += "0 ";
    if (a < D) num(j.lims[a]);
    else out += "0";
    out += "\n";
  }
  out += "ITEM: ATOMS id type x y z ";
  for (const std::string& s : j.names) {
    out += s;
    out += " ";
  }
  out += "\n";
  const int fv = D + 2, fs = 2 * D + 2;  // first velocity and stress fields
  bool ok = true;
  size_t s = 0;
  for (int64_t p = 0; p < j.np && ok; p++) {
    while (s + 1 < j.solid_end.size() && p >= j.solid_end[s]) s++;
    out += std::to_string(p + 1);
    out += " ";
    out += std::to_string(s + 1);
    out += " ";
    for (int a = 0; a < D; a++) {
      num(fld(a, p));
      out += " ";
    }
    if (D == 2) out += "0 ";
    const bool el = j.elastic[s] != 0;
    for (int c : j.cols) {
      switch (c) {
        case kColPstrain:
          if (el) out += "0 ";
          else { num(j.alpha[p]); out += " "; }
          break;
        case kColVonMises:
          if (el) out += "0 ";
          else { num(j.vm[p]); out += " "; }
          break;
        case kColSxx: num(fld(fs, p)); out += " "; break;
        case kColSyy: num(fld(fs + 1, p)); out += " "; break;
        case kColPressure:
          if (D == 2) { num(fld(fs, p) + fld(fs + 1, p)); out += " "; }
          else { num(fld(fs, p) + fld(fs + 1, p) + fld(fs + 2, p)); out += "  "; }
          break;
        case kColVx: num(fld(fv, p)); out += "  "; break;
        case kColVy: num(fld(fv + 1, p)); out += "  "; break;
        case kColVz: num(fld(fv + 2, p)); out += "  "; break;
        default: break;
      }
    }
    out += "\n";
    if (out.size() > (size_t(1) << 21)) {
      ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
      out.clear();
    }
  }
  if (ok && !out.empty()) ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
  return std::fclose(f) == 0 && ok;
}

// Column codes of a space-separated list; false for a name the path does
// not produce (damage, crackforce and color need models outside it).
inline bool parse_dump_cols(const char* outs, int dim, std::vector<std::string>& names, std::vector<int>& cols) {
  names.clear();
  cols.clear();
  if (!outs) return true;
  const char* s = outs;
  while (*s) {
    while (*s == ' ') s++;
    const char* e = s;
    while (*e && *e != ' ') e++;
    if (e == s) break;
    const std::string w(s, size_t(e - s));
    int c = -1;
    if (w == "pstrain") c = kColPstrain;
    else if (w == "vonMises") c = kColVonMises;
    else if (w == "sigmaxx" && dim == 2) c = kColSxx;
    else if (w == "sigmayy" && dim == 2) c = kColSyy;
    else if (w == "pressure") c = kColPressure;
    else if (w == "vx") c = kColVx;
    else if (w == "vy") c = kColVy;
    else if (w == "vz" && dim == 3) c = kColVz;
    if (c < 0) return false;
    names.push_back(w);
    cols.push_back(c);
    s = e;
  }
  return true;
}

// A worker thread that writes queued dumps in order.
class DumpWriter {
 public:
  ~DumpWriter() { stop(); }
  void push(DumpJob job) {
    std::unique_lock<std::mutex> lk(m_);
    if (!thread_.joinable()) {
      quit_ = false;
      thread_ = std::thread([this] { run(); });
    }
    jobs_.push_back(std::move(job));
    cv_.notify_all();
  }
  // Blocks until every queued dump is written; returns false if any failed
  // since the last call.
  bool wait() {
    std::unique_lock<std::mutex> lk(m_);
    idle_.wait(lk, [this] { return jobs_.empty() && !busy_; });
    const bool ok = !failed_;
    failed_ = false;
    return ok;
  }
  size_t pending() {
    std::unique_lock<std::mutex> lk(m_);
    return jobs_.size() + (busy_ ? 1 : 0);
  }
  void stop() {
    {
      std::unique_lock<std::mutex> lk(m_);
      quit_ = true;
      cv_.notify_all();
    }
    if (thread_.joinable()) thread_.join();
  }

 private:
  void run() {
    std::unique_lock<std::mutex> lk(m_);
    while (true) {
      cv_.wait(lk, [this] { return quit_ || !jobs_.empty(); });
      if (jobs_.empty()) break;  // quit with nothing left
      DumpJob j = std::move(jobs_.front());
      jobs_.pop_front();
      busy_ = true;
      lk.unlock();
      bool ok = j.ready ? j.ready() : true;
      ok = ok && write_dump(j);
      if (j.release) j.release();
      lk.lock();
      busy_ = false;
      if (!ok) failed_ = true;
      idle_.notify_all();
    }
    idle_.notify_all();
  }
  std::mutex m_;
  std::condition_variable cv_, idle_;
  std::deque<DumpJob> jobs_;
  std::thread thread_;
  bool quit_ = false, busy_ = false, failed_ = false;
};

}  // namespace jh
