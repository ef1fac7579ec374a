// Host build of the library's output code (jump_amd/csrc/mpm_output.hpp).
//   output_check num          read 16-hex-digit bit patterns from stdin, one
//                             per line, write Julia's text of each double
//   output_check dump IN OUT  read a small state from IN (see below) and
//                             write its LAMMPS dump to OUT
// IN (binary, little endian): int32 dim, int64 np, int32 nsolid,
//   nsolid x (int64 end, uint8 elastic), 3 doubles lims, int32 len + outs
//   text, then np*D pos, np*D vel, np*S sigma, np alpha, np vm (doubles).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../jump_amd/csrc/mpm_output.hpp"

template <class T> static T rd(std::FILE* f) {
  T v{};
  if (std::fread(&v, sizeof(T), 1, f) != 1) std::exit(3);
  return v;
}

int main(int argc, char** argv) {
  if (argc >= 2 && std::strcmp(argv[1], "num") == 0) {
    char line[64], buf[48];
    while (std::fgets(line, sizeof(line), stdin)) {
      const unsigned long long bits = std::strtoull(line, nullptr, 16);
      double x;
      std::memcpy(&x, &bits, 8);
      const int n = jh::julia_float(x, buf);
      std::fwrite(buf, 1, size_t(n), stdout);
      std::fputc('\n', stdout);
    }
    return 0;
  }
  if (argc == 4 && std::strcmp(argv[1], "dump") == 0) {
    std::FILE* f = std::fopen(argv[2], "rb");
    if (!f) return 2;
    jh::DumpJob j;
    j.dim = rd<int32_t>(f);
    j.np = rd<int64_t>(f);
    const int ns = rd<int32_t>(f);
    for (int s = 0; s < ns; s++) {
      j.solid_end.push_back(rd<int64_t>(f));
      j.elastic.push_back(rd<uint8_t>(f));
    }
    for (int a = 0; a < 3; a++) j.lims[a] = rd<double>(f);
    const int len = rd<int32_t>(f);
    std::string outs(size_t(len), ' ');
    if (len && std::fread(&outs[0], 1, size_t(len), f) != size_t(len)) return 3;
    if (!jh::parse_dump_cols(outs.c_str(), j.dim, j.names, j.cols)) return 4;
    const int D = j.dim, S = D == 3 ? 6 : 3, NF = 2 * D + 2 + S;
    const size_t n = size_t(j.np);
    std::vector<double> pos(n * D), vel(n * D), sig(n * S), alpha(n), vm(n);
    auto fill = [&](std::vector<double>& v) {
      if (!v.empty() && std::fread(v.data(), 8, v.size(), f) != v.size()) std::exit(3);
    };
    fill(pos); fill(vel); fill(sig); fill(alpha); fill(vm);
    std::fclose(f);
    // the library's pair layout (x.., m, V, v.., sigma..), m and V unused
    std::vector<std::vector<double>> pairs(size_t((NF + 1) / 2), std::vector<double>(2 * n, 0.0));
    auto put = [&](int fi, size_t p, double x) { pairs[size_t(fi / 2)][2 * p + size_t(fi % 2)] = x; };
    for (size_t p = 0; p < n; p++) {
      for (int a = 0; a < D; a++) put(a, p, pos[D * p + a]);
      for (int a = 0; a < D; a++) put(D + 2 + a, p, vel[D * p + a]);
      for (int k = 0; k < S; k++) put(2 * D + 2 + k, p, sig[S * p + k]);
    }
    for (size_t k = 0; k < pairs.size(); k++) j.pr[k] = pairs[k].data();
    j.alpha = alpha.data();
    j.vm = vm.data();
    j.path = argv[3];
    return jh::write_dump(j) ? 0 : 5;
  }
  return 1;
}
