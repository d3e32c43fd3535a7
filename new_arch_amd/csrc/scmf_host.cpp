// scmf_host.cpp — readers and writers for the reference's text formats.
//
// The readers accept what the reference's fscanf patterns accept
// (init_config.cu:216-336, 491-512): every blank in a pattern matches any run
// of whitespace, every other character must match literally.  Unlike the
// reference they report malformed input instead of reading garbage.
#include "scmf_host.h"

#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <thread>
#include <vector>

namespace scmf {

namespace {

struct Scanner {
  std::string buf;
  size_t pos = 0;
  std::string err;

  bool load(const char* path) {
    FILE* f = fopen(path, "rb");
    if (!f) {
      err = std::string("cannot open ") + path;
      return false;
    }
    char tmp[1 << 16];
    size_t n;
    while ((n = fread(tmp, 1, sizeof tmp, f)) > 0) buf.append(tmp, n);
    fclose(f);
    return true;
  }
  void skip_ws() {
    while (pos < buf.size() && isspace((unsigned char)buf[pos])) ++pos;
  }
  // fscanf literal semantics; every reference pattern ends in whitespace, so
  // leading whitespace before a literal is skipped as well.
  bool expect(const char* lit) {
    skip_ws();
    for (const char* c = lit; *c; ++c) {
      if (isspace((unsigned char)*c)) {
        skip_ws();
        continue;
      }
      if (pos >= buf.size() || buf[pos] != *c) {
        if (err.empty()) err = std::string("expected \"") + lit + "\" at byte " + std::to_string(pos);
        return false;
      }
      ++pos;
    }
    return true;
  }
  bool read_long(long long* v, const char* what) {
    skip_ws();
    const char* s = buf.c_str() + pos;
    char* end = nullptr;
    long long x = strtoll(s, &end, 10);
    if (end == s) {
      if (err.empty()) err = std::string("expected an integer for ") + what + " at byte " + std::to_string(pos);
      return false;
    }
    pos += (size_t)(end - s);
    *v = x;
    return true;
  }
  bool read_int(int32_t* v, const char* what) {
    long long x;
    if (!read_long(&x, what)) return false;
    *v = (int32_t)x;
    return true;
  }
  bool read_double(double* v, const char* what) {
    skip_ws();
    const char* s = buf.c_str() + pos;
    char* end = nullptr;
    double x = strtod(s, &end);
    if (end == s) {
      if (err.empty()) err = std::string("expected a number for ") + what + " at byte " + std::to_string(pos);
      return false;
    }
    pos += (size_t)(end - s);
    *v = x;
    return true;
  }
};

}  // namespace

scmf_desc HostDesc::view() const {
  scmf_desc d;
  d.n_mono_types = T;
  d.xn = xn.data();
  for (int i = 0; i < 3; ++i) {
    d.L[i] = L[i];
    d.n[i] = n[i];
  }
  d.dtNxi = dtNxi;
  d.reference_Nbeads = nref;
  d.n_poly_types = (int32_t)poly_length.size();
  d.poly_length = poly_length.data();
  d.n_polymers = n_polymers.data();
  d.monotype = monotype.data();
  d.connection = connection.data();
  return d;
}

std::string read_configuration(const char* path, HostDesc* d) {
  Scanner s;
  if (!s.load(path)) return s.err;
  bool ok = s.expect("Monomer type number:") && s.read_int(&d->T, "Monomer type number");
  if (!ok) return std::string(path) + ": " + s.err;
  if (d->T <= 0 || d->T > 64) return std::string(path) + ": bad monomer type number";
  // xn: file row s, column t is stored at xn[t + T*s] (init_config.cu:496-501, 670-672).
  d->xn.assign((size_t)d->T * d->T, 0.0);
  for (int r = 0; r < d->T && ok; ++r)
    for (int c = 0; c < d->T && ok; ++c) ok = s.read_double(&d->xn[c + (size_t)d->T * r], "xn");
  ok = ok && s.expect("D0 =") && s.read_double(&d->L[0], "D0") && s.expect(";") &&
       s.expect("LY =") && s.read_double(&d->L[1], "LY") && s.expect(";") && s.expect("LZ =") &&
       s.read_double(&d->L[2], "LZ") && s.expect(";");
  ok = ok && s.expect("N =") && s.read_int(&d->n[0], "nx") && s.read_int(&d->n[1], "ny") &&
       s.read_int(&d->n[2], "nz");
  ok = ok && s.expect("time =") && s.read_int(&d->time, "time") && s.expect(";");
  ok = ok && s.expect("xiN =") && s.read_double(&d->xiN, "xiN") && s.expect(";");
  ok = ok && s.expect("dt*N/xi =") && s.read_double(&d->dtNxi, "dt*N/xi") && s.expect(";");
  ok = ok && s.expect("kT =") && s.read_double(&d->kT, "kT");
  if (!ok) return std::string(path) + ": " + s.err;
  return "";
}

std::string read_polymer(const char* path, HostDesc* d) {
  Scanner s;
  if (!s.load(path)) return s.err;
  int32_t P = 0;
  long long total = 0;
  bool ok = s.expect("Number of polymer type:") && s.read_int(&P, "Number of polymer type") &&
            s.expect("Number of total polymer:") && s.read_long(&total, "Number of total polymer");
  if (!ok) return std::string(path) + ": " + s.err;
  if (P <= 0 || P > 1024) return std::string(path) + ": bad polymer type number";
  d->n_polymers.assign(P, 0);
  long long sum = 0;
  for (int p = 0; p < P && ok; ++p) {
    long long v;
    ok = s.read_long(&v, "polymer count");
    d->n_polymers[p] = v;
    sum += v;
  }
  ok = ok && s.expect("--------");
  if (!ok) return std::string(path) + ": " + s.err;
  if (sum != total)
    return std::string(path) + ": per-type polymer counts do not add up to the total";
  d->poly_length.assign(P, 0);
  d->monotype.clear();
  d->connection.clear();
  for (int p = 0; p < P; ++p) {
    int32_t idx, N;
    ok = s.expect("Polymer type:") && s.read_int(&idx, "Polymer type") &&
         s.expect("Polymer length:") && s.read_int(&N, "Polymer length") &&
         s.expect("Polymer length unit:") && s.read_int(&d->nref, "Polymer length unit") &&
         s.expect("Monomer type:");
    if (!ok) return std::string(path) + ": " + s.err;
    if (N <= 0 || N > 65536) return std::string(path) + ": bad polymer length";
    d->poly_length[p] = N;
    for (int i = 0; i < N && ok; ++i) {
      int32_t t;
      ok = s.read_int(&t, "Monomer type");
      d->monotype.push_back(t);
    }
    // Bond matrix, file row i column j (init_config.cu:291-302, 312-336).
    const size_t off = d->connection.size();
    d->connection.resize(off + (size_t)N * N, 0);
    for (int i = 0; i < N && ok; ++i)
      for (int j = 0; j < N && ok; ++j) {
        int32_t v;
        ok = s.read_int(&v, "connection");
        d->connection[off + (size_t)i * N + j] = (uint8_t)(v == 1);
      }
    if (!ok) return std::string(path) + ": " + s.err;
  }
  return "";
}

std::string validate_desc(const scmf_desc* d) {
  if (!d) return "null description";
  if (d->n_mono_types <= 0) return "n_mono_types must be positive";
  if (!d->xn) return "xn missing";
  for (int i = 0; i < 3; ++i) {
    if (!(d->L[i] > 0)) return "box lengths must be positive";
    if (d->n[i] <= 0) return "grid sizes must be positive";
  }
  if (d->reference_Nbeads < 2) return "reference_Nbeads must be at least 2";
  if (d->n_poly_types <= 0) return "n_poly_types must be positive";
  if (!d->poly_length || !d->n_polymers || !d->monotype || !d->connection)
    return "polymer arrays missing";
  int64_t chains = 0;
  size_t moff = 0, coff = 0;
  for (int p = 0; p < d->n_poly_types; ++p) {
    const int N = d->poly_length[p];
    if (N <= 0) return "poly_length must be positive";
    if (d->n_polymers[p] < 0) return "n_polymers must be non-negative";
    chains += d->n_polymers[p];
    for (int i = 0; i < N; ++i) {
      const int t = d->monotype[moff + i];
      if (t < 0 || t >= d->n_mono_types) return "monotype out of range";
      for (int j = 0; j < N; ++j) {
        const uint8_t a = d->connection[coff + (size_t)i * N + j];
        const uint8_t b = d->connection[coff + (size_t)j * N + i];
        if ((a != 0) != (b != 0)) return "bond matrix is not symmetric";
        if (i == j && a) return "bond matrix has a self bond";
      }
    }
    moff += N;
    coff += (size_t)N * N;
  }
  if (chains <= 0) return "no chains";
  if (chains > 0xFFFFFFFFll) return "more than 2^32 chains";
  return "";
}

std::string density_file_name(int32_t step) {
  char name[64];
  if (step < 1000000)
    snprintf(name, sizeof name, "density_%d.dat", step);
  else
    snprintf(name, sizeof name, "density_%ge+06.dat", (double)step / 1000000);
  return name;
}

namespace {

// printf text of a double, memoised by bit pattern in a direct-mapped table:
// the values of a density file repeat (counts are small integers, omega a
// function of them), and glibc's %g costs ~250 ns.  A hit returns the text
// snprintf produced for the same bits, so the output is unchanged.  Text
// longer than a table entry is formatted every time.
struct GFormat {
  static constexpr int BITS = 13;
  const char* fmt;
  uint64_t key[1 << BITS];
  uint8_t len[1 << BITS];  // 0: empty
  char text[1 << BITS][31];
  explicit GFormat(const char* f = "%g ") : fmt(f) { memset(len, 0, sizeof len); }
  void append(std::string& out, double v) {
    uint64_t b;
    memcpy(&b, &v, 8);
    const uint32_t h = (uint32_t)((b * 0x9E3779B97F4A7C15ull) >> (64 - BITS));
    if (len[h] == 0 || key[h] != b) {
      const int k = snprintf(text[h], sizeof text[h], fmt, v);
      if (k <= 0 || k >= (int)sizeof text[h]) {  // too long for the table
        len[h] = 0;
        std::vector<char> big((size_t)(k > 0 ? k : 0) + 1);
        snprintf(big.data(), big.size(), fmt, v);
        out.append(big.data(), big.size() - 1);
        return;
      }
      key[h] = b;
      len[h] = (uint8_t)k;
    }
    out.append(text[h], len[h]);
  }
};

// The double that the "%g" text of v reads back as (what a reader of the
// density file gets), memoised the same way.
struct GReadBack {
  static constexpr int BITS = 13;
  uint64_t key[1 << BITS];
  double val[1 << BITS];
  uint8_t used[1 << BITS];
  GReadBack() { memset(used, 0, sizeof used); }
  double get(double v) {
    uint64_t b;
    memcpy(&b, &v, 8);
    const uint32_t h = (uint32_t)((b * 0x9E3779B97F4A7C15ull) >> (64 - BITS));
    if (!used[h] || key[h] != b) {
      char t[48];
      snprintf(t, sizeof t, "%g", v);
      val[h] = strtod(t, nullptr);
      key[h] = b;
      used[h] = 1;
    }
    return val[h];
  }
};

// Formats blocks [0, blocks) of text with up to eight threads and writes
// them in order: fmt(block, out_string, worker).
template <class F>
void write_blocks(FILE* f, int64_t blocks, F fmt) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int W = (int)std::min<int64_t>(std::min(8u, hw), std::max<int64_t>(blocks, 1));
  std::vector<std::string> part((size_t)W);
  for (int64_t b0 = 0; b0 < blocks; b0 += W) {
    const int nb = (int)std::min<int64_t>(W, blocks - b0);
    auto run = [&](int w) { fmt(b0 + w, part[(size_t)w], w); };
    std::vector<std::thread> th;
    for (int w = 1; w < nb; ++w) th.emplace_back(run, w);
    run(0);
    for (auto& t : th) t.join();
    for (int w = 0; w < nb; ++w) fwrite(part[(size_t)w].data(), 1, part[(size_t)w].size(), f);
  }
}

// Closes f (written as tmp) and renames it to path.
std::string finish_file(FILE* f, const std::string& tmp, const char* path) {
  const bool bad = ferror(f) != 0;
  if (fclose(f) != 0 || bad) {
    remove(tmp.c_str());
    return std::string("write error on ") + tmp;
  }
  if (rename(tmp.c_str(), path) != 0) {
    remove(tmp.c_str());
    return std::string("cannot rename ") + tmp + " to " + path;
  }
  return "";
}

// Rows [c0, c1) of a density file: T scaled counts, then T omega values,
// each printed "%g ".
void format_density_rows(std::string& out, GFormat& g, int64_t c0, int64_t c1, int T,
                         int64_t n_cells, const uint32_t* counts, const float* omega,
                         double field_scaling) {
  out.clear();
  for (int64_t c = c0; c < c1; ++c) {
    for (int t = 0; t < T; ++t) g.append(out, counts[c + t * n_cells] * field_scaling);
    for (int t = 0; t < T; ++t) g.append(out, (double)omega[c + t * n_cells]);
    out.push_back('\n');
  }
}

}  // namespace

// io_out_density's text format (io.cpp:4-56), formatted by up to eight
// threads (each with its own GFormat) in blocks of 65,536 cells and written
// in cell order to PATH.tmp, then renamed to PATH (a reader never sees a
// partial file).
std::string write_density_file(const char* path, const int32_t n[3], const double L[3], int T,
                               int64_t n_cells, const uint32_t* counts, const float* omega,
                               double field_scaling) {
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = fopen(tmp.c_str(), "w");
  if (!f) return std::string("cannot open ") + tmp + " for writing";
  fprintf(f, "Nx=%d, Ny=%d, Nz=%d\n", n[0], n[1], n[2]);
  fprintf(f, "dx=%g, dy=%g, dz=%g\n", L[0] / n[0], L[1] / n[1], L[2] / n[2]);
  constexpr int64_t ROWS = 1 << 16;
  std::vector<std::unique_ptr<GFormat>> cache(8);
  for (auto& g : cache) g.reset(new GFormat());
  write_blocks(f, (n_cells + ROWS - 1) / ROWS, [&](int64_t blk, std::string& out, int w) {
    const int64_t c0 = blk * ROWS;
    format_density_rows(out, *cache[(size_t)w], c0, std::min(c0 + ROWS, n_cells), T, n_cells,
                        counts, omega, field_scaling);
  });
  return finish_file(f, tmp, path);
}

// The VTK file density_to_vtk makes of that density file, straight from the
// counts and omega: every value goes through its "%g" text and back, as the
// converter reads it, and each point is the first column minus the second
// (type 0 minus type 1, or minus omega_0 when T = 1).  density_name is the
// name the converter's comment line carries.
std::string write_vtk_file(const char* path, const char* density_name, const int32_t n[3],
                           const double L[3], int T, int64_t n_cells, const uint32_t* counts,
                           const float* omega, double field_scaling) {
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = fopen(tmp.c_str(), "w");
  if (!f) return std::string("cannot open ") + tmp + " for writing";
  double d[3];
  for (int a = 0; a < 3; ++a) {
    char t[48];
    snprintf(t, sizeof t, "%g", L[a] / n[a]);
    d[a] = strtod(t, nullptr);
  }
  fprintf(f, "# vtk DataFile Version 3.0\n");
  fprintf(f, "# comment: converted from file %s\n", density_name);
  fprintf(f, "ASCII\n");
  fprintf(f, "DATASET STRUCTURED_POINTS\n");
  fprintf(f, "DIMENSIONS %d %d %d\n", n[0], n[1], n[2]);
  fprintf(f, "ORIGIN 0.000000 0.000000 0.000000\n");
  fprintf(f, "SPACING %.6lf %.6lf %.6lf\n\n", d[0], d[1], d[2]);
  fprintf(f, "POINT_DATA %ld\n", (long)n[0] * n[1] * n[2]);
  fprintf(f, "SCALARS scalars float\n");
  fprintf(f, "LOOKUP_TABLE default\n\n");
  const int64_t rows = (int64_t)n[1] * n[2];
  const int64_t per = std::max<int64_t>(1, (1 << 16) / n[0]);
  std::vector<std::unique_ptr<GFormat>> fmt(8);
  std::vector<std::unique_ptr<GReadBack>> rb(8);
  for (int w = 0; w < 8; ++w) {
    fmt[(size_t)w].reset(new GFormat("%7.4lf "));
    rb[(size_t)w].reset(new GReadBack());
  }
  write_blocks(f, (rows + per - 1) / per, [&](int64_t blk, std::string& out, int w) {
    out.clear();
    GFormat& g = *fmt[(size_t)w];
    GReadBack& r = *rb[(size_t)w];
    const int64_t r1 = std::min(rows, (blk + 1) * per);
    for (int64_t row = blk * per; row < r1; ++row) {
      for (int i = 0; i < n[0]; ++i) {
        const int64_t c = row * n[0] + i;
        const double a = r.get(counts[c] * field_scaling);
        const double b = T >= 2 ? r.get(counts[c + n_cells] * field_scaling) : r.get((double)omega[c]);
        g.append(out, a - b);
      }
      out.push_back('\n');
    }
  });
  return finish_file(f, tmp, path);
}

std::string write_coord_file(const char* path, const float* xyz, int64_t n_beads) {
  FILE* f = fopen(path, "w");
  if (!f) return std::string("cannot open ") + path + " for writing";
  std::unique_ptr<char[]> big(new char[1 << 20]);
  setvbuf(f, big.get(), _IOFBF, 1 << 20);
  for (int64_t b = 0; b < n_beads; ++b)
    fprintf(f, "%g %g %g\n", (double)xyz[3 * b], (double)xyz[3 * b + 1], (double)xyz[3 * b + 2]);
  const bool bad = ferror(f) != 0;
  fclose(f);
  return bad ? std::string("write error on ") + path : "";
}

std::string density_to_vtk(const char* in, const char* out) {
  Scanner s;
  if (!s.load(in)) return s.err;
  int32_t nx, ny, nz;
  double dx, dy, dz;
  bool ok = s.expect("Nx=") && s.read_int(&nx, "Nx") && s.expect(", Ny=") &&
            s.read_int(&ny, "Ny") && s.expect(", Nz=") && s.read_int(&nz, "Nz") &&
            s.expect("dx=") && s.read_double(&dx, "dx") && s.expect(", dy=") &&
            s.read_double(&dy, "dy") && s.expect(", dz=") && s.read_double(&dz, "dz");
  if (!ok) return std::string(in) + ": " + s.err;
  if (nx <= 0 || ny <= 0 || nz <= 0) return std::string(in) + ": bad grid";
  // Columns per row: count the numbers on the first data line.
  s.skip_ws();
  size_t eol = s.buf.find('\n', s.pos);
  if (eol == std::string::npos) eol = s.buf.size();
  int cols = 0;
  {
    const char* p = s.buf.c_str() + s.pos;
    const char* e = s.buf.c_str() + eol;
    while (p < e) {
      char* q;
      strtod(p, &q);
      if (q == p) break;
      ++cols;
      p = q;
      while (p < e && isspace((unsigned char)*p)) ++p;
    }
  }
  if (cols < 2) return std::string(in) + ": rows need at least two columns";
  FILE* f = fopen(out, "w");
  if (!f) return std::string("cannot open ") + out + " for writing";
  fprintf(f, "# vtk DataFile Version 3.0\n");
  fprintf(f, "# comment: converted from file %s\n", in);
  fprintf(f, "ASCII\n");
  fprintf(f, "DATASET STRUCTURED_POINTS\n");
  fprintf(f, "DIMENSIONS %d %d %d\n", nx, ny, nz);
  fprintf(f, "ORIGIN 0.000000 0.000000 0.000000\n");
  fprintf(f, "SPACING %.6lf %.6lf %.6lf\n\n", dx, dy, dz);
  fprintf(f, "POINT_DATA %ld\n", (long)nx * ny * nz);
  fprintf(f, "SCALARS scalars float\n");
  fprintf(f, "LOOKUP_TABLE default\n\n");
  std::vector<double> row(cols);
  for (int k = 0; k < nz && ok; ++k) {
    for (int j = 0; j < ny && ok; ++j) {
      for (int i = 0; i < nx && ok; ++i) {
        for (int c = 0; c < cols && ok; ++c) ok = s.read_double(&row[c], "density row");
        if (ok) fprintf(f, "%7.4lf ", row[0] - row[1]);
      }
      fprintf(f, "\n");
    }
  }
  fclose(f);
  if (!ok) return std::string(in) + ": " + s.err;
  return "";
}

}  // namespace scmf
