// io.cpp -- native readers and writers of the on-disk formats: data files
// (read_data) and the atom and element dumps.  Host code only.
//
// A data file is read whole, indexed by line, and each large section
// (Elements, Nodes, Atoms) is parsed in parallel chunks of lines; the
// writers format chunks of rows in parallel into separate buffers and write
// them in order.  Numbers are parsed the way the reference's atoi / atof
// do (the leading number of a word, 0 if there is none) and formatted with
// std::to_chars, which prints exactly what printf's %g / %.17g print.
#include <algorithm>
#include <atomic>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "cac_io.h"

namespace {

thread_local std::string g_err;

int io_fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

struct DataError {
  std::string msg;
};

// ------------------------------------------------------------ parallelism
int io_threads() {
  if (const char *e = std::getenv("CAC_IO_THREADS")) {
    const int v = std::atoi(e);
    if (v > 0) return v;
  }
  const unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(h ? h : 1u, 16u));
}

// fn(begin, end, chunk) over [0, n) in `chunks` contiguous ranges, each on
// its own thread (the calling thread takes the first)
void parallel_chunks(int64_t n, int chunks, const std::function<void(int64_t, int64_t, int)> &fn) {
  if (chunks <= 1 || n < 4096) {
    fn(0, n, 0);
    return;
  }
  std::vector<std::thread> th;
  const int64_t step = (n + chunks - 1) / chunks;
  for (int c = 1; c < chunks; ++c) {
    const int64_t b = std::min(n, c * step), e = std::min(n, b + step);
    th.emplace_back(fn, b, e, c);
  }
  fn(0, std::min(n, step), 0);
  for (auto &t : th) t.join();
}

// ------------------------------------------------------------ numbers
// atoi / atoll: the leading integer of a word, 0 if it has none
int64_t word_int(const char *b, const char *e) {
  if (b < e && *b == '+') ++b;
  int64_t v = 0;
  const auto r = std::from_chars(b, e, v);
  return r.ec == std::errc() ? v : 0;
}
// atof: the leading number of a word, 0 if it has none
double word_double(const char *b, const char *e) {
  if (b < e && *b == '+') ++b;
  double v = 0.0;
  const auto r = std::from_chars(b, e, v);
  return r.ec == std::errc() ? v : 0.0;
}

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\f' || c == '\v'; }

// words of [b, e): up to maxw (begin, end) pairs, returns the word count
int split_words(const char *b, const char *e, const char **wb, const char **we, int maxw) {
  int n = 0;
  while (b < e) {
    while (b < e && is_space(*b)) ++b;
    if (b >= e) break;
    const char *s = b;
    while (b < e && !is_space(*b)) ++b;
    if (n < maxw) {
      wb[n] = s;
      we[n] = b;
    }
    ++n;
  }
  return n;
}

// ------------------------------------------------------------ the file
struct Lines {
  std::string buf;
  std::vector<size_t> start;  // line i is [start[i], start[i+1] - 1)
  int64_t count() const { return (int64_t)start.size() - 1; }
  const char *b(int64_t i) const { return buf.data() + start[i]; }
  const char *e(int64_t i) const { return buf.data() + start[i + 1] - 1; }
  std::string str(int64_t i) const { return std::string(b(i), e(i)); }
};

// read a whole file and index its lines (split at '\n'; a final newline ends
// the last line rather than starting an empty one)
int load_lines(const char *path, Lines &L) {
  FILE *fp = std::fopen(path, "rb");
  if (!fp) return io_fail(CAC_ERR_IO, std::string("Cannot open file ") + path);
  std::fseek(fp, 0, SEEK_END);
  const long sz = std::ftell(fp);
  std::fseek(fp, 0, SEEK_SET);
  if (sz < 0) {
    std::fclose(fp);
    return io_fail(CAC_ERR_IO, std::string("Cannot read file ") + path);
  }
  L.buf.resize((size_t)sz + 1);
  const size_t got = sz ? std::fread(&L.buf[0], 1, (size_t)sz, fp) : 0;
  std::fclose(fp);
  if (got != (size_t)sz) return io_fail(CAC_ERR_IO, std::string("Cannot read file ") + path);
  size_t len = (size_t)sz;
  if (len && L.buf[len - 1] == '\n') --len;  // the newline that ends the last line
  L.buf[len] = '\n';
  L.buf.resize(len + 1);
  L.start.clear();
  L.start.push_back(0);
  const char *p = L.buf.data(), *end = p + L.buf.size();
  while (p < end) {
    const char *q = static_cast<const char *>(std::memchr(p, '\n', (size_t)(end - p)));
    L.start.push_back((size_t)(q - L.buf.data()) + 1);
    p = q + 1;
  }
  return CAC_OK;
}

// text before a '#', trimmed
std::string uncommented(const char *b, const char *e) {
  const char *h = static_cast<const char *>(std::memchr(b, '#', (size_t)(e - b)));
  if (h) e = h;
  while (b < e && is_space(*b)) ++b;
  while (e > b && is_space(e[-1])) --e;
  return std::string(b, e);
}

}  // namespace

struct cac_data {
  cac_data_view view;
  std::string comment;
  std::vector<double> masses;
  std::vector<int> nae, numint, intp_ijk;
  std::vector<double> intp_w;
  std::vector<int64_t> atom_tag, elem_tag, nodetag;
  std::vector<int> atom_type, atom_image, elem_type, elem_image;
  std::vector<double> x, xc, nodex, nodev;
  bool has_masses = false, has_nae = false, has_numint = false, has_intp = false;
};

namespace {

// ------------------------------------------------------------ data file parse
// Parse `rows` lines from line i0 of a section whose lines all have the
// word count of the first (nw, one of the allowed two), in parallel chunks.
// row(r, words...) stores one line.  A line with another word count is
// `what`'s format error.
void parse_rows(const Lines &L, int64_t i0, int64_t rows, int nw, const char *what,
                const std::function<void(int64_t, const char **, const char **)> &row) {
  std::vector<std::string> errs(64);
  const int chunks = std::min(io_threads(), 64);
  parallel_chunks(rows, chunks, [&](int64_t b, int64_t e, int c) {
    const char *wb[16], *we[16];
    for (int64_t r = b; r < e; ++r) {
      const int n = split_words(L.b(i0 + r), L.e(i0 + r), wb, we, 16);
      if (n != nw) {
        errs[c] = std::string("Incorrect ") + what + " format in data file";
        return;
      }
      row(r, wb, we);
    }
  });
  for (auto &m : errs)
    if (!m.empty()) throw DataError{m};
}

int count_words(const Lines &L, int64_t i) {
  const char *wb[1], *we[1];
  return split_words(L.b(i), L.e(i), wb, we, 0);
}

void parse_file(const Lines &L, cac_data &d) {
  const int64_t nl = L.count();
  if (nl == 0 || (nl == 1 && L.b(0) == L.e(0))) throw DataError{"Unexpected end of data file"};
  d.comment = L.str(0);
  if (!d.comment.empty() && d.comment.back() == '\r') d.comment.pop_back();
  cac_data_view &v = d.view;
  int64_t natoms = 0, nelements = 0;
  v.ntypes = v.entypes = 1;
  for (int c = 0; c < 3; ++c) v.lo[c] = v.hi[c] = 0.0;
  int64_t i = 1;
  std::string keyword;
  // header (ReadData::header, read_data.cpp:277-368): the first non-blank
  // line without a header keyword is the first section keyword
  auto first_int = [](const std::string &s) {
    const char *wb[1], *we[1];
    return split_words(s.data(), s.data() + s.size(), wb, we, 1) ? word_int(wb[0], we[0]) : 0;
  };
  auto two_doubles = [](const std::string &s, double *a, double *b) {
    const char *wb[2], *we[2];
    const int n = split_words(s.data(), s.data() + s.size(), wb, we, 2);
    *a = n > 0 ? word_double(wb[0], we[0]) : 0.0;
    *b = n > 1 ? word_double(wb[1], we[1]) : 0.0;
  };
  while (i < nl) {
    const std::string line = uncommented(L.b(i), L.e(i));
    ++i;
    if (line.empty()) continue;
    if (line.find("atoms") != std::string::npos)
      natoms = first_int(line);
    else if (line.find("elements") != std::string::npos)
      nelements = first_int(line);
    else if (line.find("xlo xhi") != std::string::npos)
      two_doubles(line, &v.lo[0], &v.hi[0]);
    else if (line.find("ylo yhi") != std::string::npos)
      two_doubles(line, &v.lo[1], &v.hi[1]);
    else if (line.find("zlo zhi") != std::string::npos)
      two_doubles(line, &v.lo[2], &v.hi[2]);
    else if (line.find("atom types") != std::string::npos)
      v.ntypes = (int)first_int(line);
    else if (line.find("element types") != std::string::npos)
      v.entypes = (int)first_int(line);
    else {
      keyword = line;
      break;
    }
  }
  if (natoms < 0 || nelements < 0) throw DataError{"System in data file is too big"};
  if (v.ntypes < 1 || v.entypes < 1) throw DataError{"Invalid type count in data file"};
  auto take = [&](int64_t k) {
    if (i + k > nl) throw DataError{"Unexpected end of data file"};
    const int64_t at = i;
    i += k;
    return at;
  };
  d.masses.assign(v.ntypes, std::numeric_limits<double>::quiet_NaN());
  d.nae.assign(3 * (size_t)v.entypes, 0);
  d.numint.assign(v.entypes, -1);
  std::unordered_map<int64_t, int64_t> elem_row;
  bool have_atoms = false, have_elems = false;
  static const char *kSections[] = {"Atoms", "Masses", "NAAE", "Numint", "Intp", "Elements", "Nodes"};
  while (!keyword.empty()) {
    if (std::none_of(std::begin(kSections), std::end(kSections), [&](const char *s) { return keyword == s; }))
      throw DataError{"Unknown identifier in data file: " + keyword};
    take(1);  // the line after the keyword (read_data.cpp:397)
    if (keyword == "Masses") {
      const int64_t at = take(v.ntypes);
      for (int64_t r = 0; r < v.ntypes; ++r) {
        const char *wb[2], *we[2];
        const int n = split_words(L.b(at + r), L.e(at + r), wb, we, 2);
        const int64_t t = n > 0 ? word_int(wb[0], we[0]) : 0;
        if (t < 1 || t > v.ntypes) throw DataError{"Invalid type for mass set"};
        d.masses[t - 1] = n > 1 ? word_double(wb[1], we[1]) : 0.0;
      }
      d.has_masses = true;
    } else if (keyword == "NAAE") {  // Element::set_nae, element.cpp:1086-1115
      const int64_t at = take(v.entypes);
      for (int64_t r = 0; r < v.entypes; ++r) {
        const char *wb[4], *we[4];
        if (split_words(L.b(at + r), L.e(at + r), wb, we, 4) < 4) throw DataError{"Invalid nae line in data file"};
        const int64_t t = word_int(wb[0], we[0]);
        if (t < 1 || t > v.entypes) throw DataError{"Invalid type for nae set"};
        int nv[3];
        for (int c = 0; c < 3; ++c) nv[c] = (int)word_int(wb[1 + c], we[1 + c]);
        if (nv[0] < 2 || nv[1] < 2 || nv[2] < 2) throw DataError{"Invalid naae values"};
        for (int c = 0; c < 3; ++c) d.nae[3 * (t - 1) + c] = nv[c];
      }
      d.has_nae = true;
    } else if (keyword == "Numint") {  // Element::set_numint, element.cpp:1126-1140
      const int64_t at = take(v.entypes);
      for (int64_t r = 0; r < v.entypes; ++r) {
        const char *wb[2], *we[2];
        if (split_words(L.b(at + r), L.e(at + r), wb, we, 2) < 2) throw DataError{"Invalid numint line in data file"};
        const int64_t t = word_int(wb[0], we[0]);
        if (t < 1 || t > v.entypes) throw DataError{"Invalid type for numint set"};
        d.numint[t - 1] = (int)word_int(wb[1], we[1]);
      }
      d.has_numint = true;
    } else if (keyword == "Intp") {  // Element::set_intp, element.cpp:1144-1166
      std::vector<int64_t> first(v.entypes + 1, 0);
      for (int t = 0; t < v.entypes; ++t) first[t + 1] = first[t] + std::max(0, d.numint[t]);
      const int64_t total = first[v.entypes];
      const int64_t at = take(total);
      d.intp_ijk.assign(3 * (size_t)total, 0);
      d.intp_w.assign((size_t)total, 0.0);
      for (int64_t r = 0; r < total; ++r) {
        const char *wb[6], *we[6];
        const int n = split_words(L.b(at + r), L.e(at + r), wb, we, 6);
        if (n < 6) throw DataError{"Invalid Intp line in data file"};
        const int64_t t = word_int(wb[0], we[0]), m = word_int(wb[1], we[1]) - 1;
        if (t < 1 || t > v.entypes || d.numint[t - 1] < 0 || m < 0 || m >= d.numint[t - 1])
          throw DataError{"Invalid Intp line in data file"};
        const int64_t row = first[t - 1] + m;
        for (int c = 0; c < 3; ++c) d.intp_ijk[3 * row + c] = (int)word_int(wb[2 + c], we[2 + c]);
        d.intp_w[row] = word_double(wb[5], we[5]);
      }
      d.has_intp = true;
    } else if (keyword == "Elements") {  // Element::data_elements, element.cpp:195-314
      const int64_t at = take(nelements);
      const int nw = nelements ? count_words(L, at) : 5;
      if (nw != 5 && nw != 8) throw DataError{"Incorrect element format in data file"};
      d.elem_tag.assign(nelements, 0);
      d.elem_type.assign(nelements, 0);
      d.xc.assign(3 * (size_t)nelements, 0.0);
      if (nw == 8) d.elem_image.assign(3 * (size_t)nelements, 0);
      std::atomic<int> bad{0};
      const int et = v.entypes;
      parse_rows(L, at, nelements, nw, "element", [&](int64_t r, const char **wb, const char **we) {
        d.elem_tag[r] = word_int(wb[0], we[0]);
        const int64_t t = word_int(wb[1], we[1]);
        if (t < 1 || t > et) bad.store(1, std::memory_order_relaxed);
        d.elem_type[r] = (int)t;
        for (int c = 0; c < 3; ++c) d.xc[3 * r + c] = word_double(wb[2 + c], we[2 + c]);
        if (nw == 8)
          for (int c = 0; c < 3; ++c) d.elem_image[3 * r + c] = (int)word_int(wb[5 + c], we[5 + c]);
      });
      if (bad.load()) throw DataError{"Invalid atom type in Elements section of data file"};
      elem_row.reserve((size_t)nelements * 2);
      for (int64_t r = 0; r < nelements; ++r) elem_row.emplace(d.elem_tag[r], r);
      d.nodex.assign(24 * (size_t)nelements, 0.0);
      d.nodev.assign(24 * (size_t)nelements, 0.0);
      d.nodetag.assign(8 * (size_t)nelements, 0);
      have_elems = true;
    } else if (keyword == "Nodes") {  // Element::data_nodes, element.cpp:478-540
      const int64_t rows = 8 * nelements;
      const int64_t at = take(rows);
      if (!have_elems && nelements) throw DataError{"Nodes section before Elements in data file"};
      const int nw = rows ? count_words(L, at) : 6;
      if (nw != 6 && nw != 9) throw DataError{"Incorrect node format in data file"};
      std::atomic<int> badk{0};
      parse_rows(L, at, rows, nw, "node", [&](int64_t r, const char **wb, const char **we) {
        const auto it = elem_row.find(word_int(wb[0], we[0]));
        if (it == elem_row.end()) return;  // nodes of unknown elements are skipped (element.cpp:496)
        const int64_t k = word_int(wb[1], we[1]) - 1;
        if (k < 0 || k > 7) {
          badk.store(1, std::memory_order_relaxed);
          return;
        }
        const int64_t s = 8 * it->second + k;
        for (int c = 0; c < 3; ++c) d.nodex[3 * s + c] = word_double(wb[2 + c], we[2 + c]);
        if (nw == 9)
          for (int c = 0; c < 3; ++c) d.nodev[3 * s + c] = word_double(wb[5 + c], we[5 + c]);
        d.nodetag[s] = word_int(wb[nw - 1], we[nw - 1]);
      });
      if (badk.load()) throw DataError{"Invalid node index in Nodes section of data file"};
      v.node_velocities = nw == 9;
    } else if (keyword == "Atoms") {  // AtomVecAtomic::data_atom, atom_vec_atomic.cpp:49-71
      const int64_t at = take(natoms);
      const int nw = natoms ? count_words(L, at) : 5;
      if (nw != 5 && nw != 8) throw DataError{"Incorrect atom format in data file"};
      d.atom_tag.assign(natoms, 0);
      d.atom_type.assign(natoms, 0);
      d.x.assign(3 * (size_t)natoms, 0.0);
      if (nw == 8) d.atom_image.assign(3 * (size_t)natoms, 0);
      std::atomic<int> bad{0};
      const int nt = v.ntypes;
      parse_rows(L, at, natoms, nw, "atom", [&](int64_t r, const char **wb, const char **we) {
        d.atom_tag[r] = word_int(wb[0], we[0]);
        const int64_t t = word_int(wb[1], we[1]);
        if (t < 1 || t > nt) bad.store(1, std::memory_order_relaxed);
        d.atom_type[r] = (int)t;
        for (int c = 0; c < 3; ++c) d.x[3 * r + c] = word_double(wb[2 + c], we[2 + c]);
        if (nw == 8)
          for (int c = 0; c < 3; ++c) d.atom_image[3 * r + c] = (int)word_int(wb[5 + c], we[5 + c]);
      });
      if (bad.load()) throw DataError{"Invalid atom type in Atoms section of data file"};
      have_atoms = true;
    }
    // next keyword (ReadData::parse_keyword): the next non-blank line that
    // is not a comment
    keyword.clear();
    while (i < nl) {
      const char *b = L.b(i), *e = L.e(i);
      ++i;
      while (b < e && is_space(*b)) ++b;
      if (b == e || *b == '#') continue;
      keyword = uncommented(b, e);
      break;
    }
  }
  if (natoms > 0 && !have_atoms) throw DataError{"No atoms in data file"};
  if (nelements > 0 && !have_elems) throw DataError{"No elements in data file"};
  if (!have_atoms) natoms = 0;
  if (!have_elems) nelements = 0;
  v.natoms = natoms;
  v.nelements = nelements;
}

void fill_view(cac_data &d) {
  cac_data_view &v = d.view;
  v.comment = d.comment.c_str();
  v.masses = d.has_masses ? d.masses.data() : nullptr;
  v.nae = d.has_nae ? d.nae.data() : nullptr;
  v.numint = d.has_numint ? d.numint.data() : nullptr;
  v.nintp = d.has_intp ? (int64_t)d.intp_w.size() : 0;
  v.intp_ijk = d.has_intp ? d.intp_ijk.data() : nullptr;
  v.intp_w = d.has_intp ? d.intp_w.data() : nullptr;
  v.atom_tag = d.atom_tag.data();
  v.atom_type = d.atom_type.data();
  v.x = d.x.data();
  v.atom_image = d.atom_image.empty() ? nullptr : d.atom_image.data();
  v.elem_tag = d.elem_tag.data();
  v.elem_type = d.elem_type.data();
  v.xc = d.xc.data();
  v.elem_image = d.elem_image.empty() ? nullptr : d.elem_image.data();
  v.nodex = d.nodex.data();
  v.nodev = d.nodev.data();
  v.nodetag = d.nodetag.data();
}

// ------------------------------------------------------------ formatting
struct Out {
  std::string s;
  void g17(double v) { num(v, 17); }
  void g6(double v) { num(v, 6); }
  void num(double v, int prec) {
    char b[40];
    const auto r = std::to_chars(b, b + sizeof b, v, std::chars_format::general, prec);
    s.append(b, r.ptr);
  }
  void i(int64_t v) {
    char b[24];
    const auto r = std::to_chars(b, b + sizeof b, v);
    s.append(b, r.ptr);
  }
  void c(char ch) { s.push_back(ch); }
  void str(const char *t) { s.append(t); }
};

// rows [0, n) formatted by fn(out, row) in parallel chunks, appended to f in
// order
bool write_rows(FILE *f, int64_t n, const std::function<void(Out &, int64_t)> &fn) {
  const int chunks = std::min<int64_t>(io_threads(), std::max<int64_t>(1, n / 4096));
  std::vector<Out> part(std::max(chunks, 1));
  parallel_chunks(n, chunks, [&](int64_t b, int64_t e, int c) {
    Out &o = part[c];
    o.s.reserve((size_t)(e - b) * 96);
    for (int64_t r = b; r < e; ++r) fn(o, r);
  });
  for (auto &o : part)
    if (!o.s.empty() && std::fwrite(o.s.data(), 1, o.s.size(), f) != o.s.size()) return false;
  return true;
}

std::string step_path(const char *path, int64_t step) {
  std::string p(path);
  const size_t star = p.find('*');
  if (star != std::string::npos) p = p.substr(0, star) + std::to_string(step) + p.substr(star + 1);
  return p;
}

}  // namespace

extern "C" const char *cac_io_last_error(void) { return g_err.c_str(); }

extern "C" int cac_data_read(const char *path, cac_data **out) {
  if (!path || !out) return io_fail(CAC_ERR_ARG, "cac_data_read: null argument");
  *out = nullptr;
  Lines L;
  int rc = load_lines(path, L);
  if (rc) return rc;
  cac_data *d = new cac_data();
  std::memset(&d->view, 0, sizeof d->view);
  try {
    parse_file(L, *d);
  } catch (const DataError &e) {
    delete d;
    return io_fail(CAC_ERR_DATA, e.msg);
  }
  fill_view(*d);
  *out = d;
  return CAC_OK;
}

extern "C" const cac_data_view *cac_data_get(const cac_data *d) { return d ? &d->view : nullptr; }

extern "C" void cac_data_free(cac_data *d) { delete d; }

extern "C" int cac_data_write(const char *path, const cac_data_view *v, int node_velocities) {
  if (!path || !v) return io_fail(CAC_ERR_ARG, "cac_data_write: null argument");
  const int64_t N = v->natoms, M = v->nelements;
  if (N < 0 || M < 0 || (N && (!v->atom_tag || !v->atom_type || !v->x)) ||
      (M && (!v->elem_tag || !v->elem_type || !v->xc || !v->nodex || !v->nodetag || !v->nae || !v->numint ||
             (v->nintp && (!v->intp_ijk || !v->intp_w)) || (node_velocities && !v->nodev))))
    return io_fail(CAC_ERR_ARG, "cac_data_write: missing arrays");
  FILE *f = std::fopen(path, "wb");
  if (!f) return io_fail(CAC_ERR_IO, std::string("Cannot open file ") + path);
  Out h;
  std::string comment = v->comment ? v->comment : "";
  std::replace(comment.begin(), comment.end(), '\n', ' ');
  h.str(comment.c_str());
  h.str("\n\n");
  h.i(N);
  h.str(" atoms\n");
  h.i(M);
  h.str(" elements\n");
  h.i(v->ntypes);
  h.str(" atom types\n");
  h.i(v->entypes);
  h.str(" element types\n");
  static const char *kAx[3] = {" xlo xhi\n", " ylo yhi\n", " zlo zhi\n"};
  for (int c = 0; c < 3; ++c) {
    h.g17(v->lo[c]);
    h.c(' ');
    h.g17(v->hi[c]);
    h.str(kAx[c]);
  }
  // every header and section line ends with a newline, and one blank line
  // separates the header and the sections (none after the last)
  bool ok = true, blank = true;
  auto flush = [&](Out &o) {
    if (!o.s.empty() && std::fwrite(o.s.data(), 1, o.s.size(), f) != o.s.size()) ok = false;
    o.s.clear();
  };
  auto begin = [&](const char *name) {
    if (blank) h.c('\n');
    h.str(name);
    h.str("\n\n");  // the keyword and the line after it (read_data.cpp:397)
    blank = true;
  };
  if (v->masses) {
    bool any = false;
    for (int t = 0; t < v->ntypes; ++t) any |= !std::isnan(v->masses[t]);
    if (any) {
      begin("Masses");
      for (int t = 0; t < v->ntypes; ++t)
        if (!std::isnan(v->masses[t])) {
          h.i(t + 1);
          h.c(' ');
          h.g17(v->masses[t]);
          h.c('\n');
        }
    }
  }
  if (M) {
    begin("NAAE");
    for (int t = 0; t < v->entypes; ++t)
      if (v->nae[3 * t] > 0) {
        h.i(t + 1);
        for (int c = 0; c < 3; ++c) {
          h.c(' ');
          h.i(v->nae[3 * t + c]);
        }
        h.c('\n');
      }
    begin("Numint");
    for (int t = 0; t < v->entypes; ++t)
      if (v->numint[t] >= 0) {
        h.i(t + 1);
        h.c(' ');
        h.i(v->numint[t]);
        h.c('\n');
      }
    begin("Intp");
    int64_t row = 0;
    for (int t = 0; t < v->entypes; ++t)
      for (int m = 0; m < std::max(0, v->numint[t]) && row < v->nintp; ++m, ++row) {
        h.i(t + 1);
        h.c(' ');
        h.i(m + 1);
        for (int c = 0; c < 3; ++c) {
          h.c(' ');
          h.i(v->intp_ijk[3 * row + c]);
        }
        h.c(' ');
        h.g17(v->intp_w[row]);
        h.c('\n');
      }
    begin("Elements");
    flush(h);
    ok = ok && write_rows(f, M, [&](Out &o, int64_t e) {
      o.i(v->elem_tag[e]);
      o.c(' ');
      o.i(v->elem_type[e]);
      for (int c = 0; c < 3; ++c) {
        o.c(' ');
        o.g17(v->xc[3 * e + c]);
      }
      if (v->elem_image)
        for (int c = 0; c < 3; ++c) {
          o.c(' ');
          o.i(v->elem_image[3 * e + c]);
        }
      o.c('\n');
    });
    begin("Nodes");
    flush(h);
    ok = ok && write_rows(f, 8 * M, [&](Out &o, int64_t s) {
      o.i(v->elem_tag[s / 8]);
      o.c(' ');
      o.i(s % 8 + 1);
      for (int c = 0; c < 3; ++c) {
        o.c(' ');
        o.g17(v->nodex[3 * s + c]);
      }
      if (node_velocities)
        for (int c = 0; c < 3; ++c) {
          o.c(' ');
          o.g17(v->nodev[3 * s + c]);
        }
      o.c(' ');
      o.i(v->nodetag[s]);
      o.c('\n');
    });
  }
  if (N) {
    begin("Atoms");
    flush(h);
    ok = ok && write_rows(f, N, [&](Out &o, int64_t i) {
      o.i(v->atom_tag[i]);
      o.c(' ');
      o.i(v->atom_type[i]);
      for (int c = 0; c < 3; ++c) {
        o.c(' ');
        o.g17(v->x[3 * i + c]);
      }
      if (v->atom_image)
        for (int c = 0; c < 3; ++c) {
          o.c(' ');
          o.i(v->atom_image[3 * i + c]);
        }
      o.c('\n');
    });
  }
  flush(h);
  if (std::fclose(f) != 0) ok = false;
  return ok ? CAC_OK : io_fail(CAC_ERR_IO, std::string("Cannot write file ") + path);
}

extern "C" int cac_dump_atom_write(const char *path, int append, int64_t step, const double lo[3],
                                   const double hi[3], const int periodic[3], int64_t n, const int64_t *tag,
                                   const int *type, const double *x, const double *f, const double *v, int64_t ne,
                                   const int64_t *nodetag, const double *nodex, const double *nodef,
                                   const double *nodev) {
  if (!path || !lo || !hi || !periodic || n < 0 || ne < 0 || (n && (!x || !f || !v)) ||
      (ne && (!nodex || !nodef || !nodev)))
    return io_fail(CAC_ERR_ARG, "cac_dump_atom_write: bad arguments");
  const std::string p = step_path(path, step);
  FILE *fp = std::fopen(p.c_str(), append ? "ab" : "wb");
  if (!fp) return io_fail(CAC_ERR_IO, "Cannot open dump file " + p);
  // DumpAtom::header_item (dump_atom.cpp:76-87); Domain::boundary_string
  // (domain.cpp:849-862): two letters per dimension, space separated
  Out h;
  h.str("ITEM: TIMESTEP\n");
  h.i(step);
  h.str("\nITEM: NUMBER OF ATOMS\n");
  h.i(n + 8 * ne);
  h.str("\nITEM: BOX BOUNDS ");
  for (int d = 0; d < 3; ++d) {
    const char b = periodic[d] ? 'p' : 'f';
    h.c(b);
    h.c(b);
    if (d < 2) h.c(' ');
  }
  h.c('\n');
  for (int d = 0; d < 3; ++d) {
    h.g6(lo[d]);
    h.c(' ');
    h.g6(hi[d]);
    h.c('\n');
  }
  h.str("ITEM: ATOMS id type x y z fx fy fz vx vy vz\n");
  bool ok = std::fwrite(h.s.data(), 1, h.s.size(), fp) == h.s.size();
  // convert_noimage (dump_atom.cpp:158-182): "%d %d %g %g %g %g %g %g %g %g %g"
  auto line = [](Out &o, int64_t id, int t, const double *a, const double *b, const double *c) {
    o.i(id);
    o.c(' ');
    o.i(t);
    for (const double *q : {a, b, c})
      for (int k = 0; k < 3; ++k) {
        o.c(' ');
        o.g6(q[k]);
      }
    o.c('\n');
  };
  ok = ok && write_rows(fp, n, [&](Out &o, int64_t i) {
    line(o, tag ? tag[i] : i + 1, type ? type[i] : 1, x + 3 * i, f + 3 * i, v + 3 * i);
  });
  // element nodes as type-1 rows (DumpAtom::epack_noscale_noimage, :121-150)
  ok = ok && write_rows(fp, 8 * ne, [&](Out &o, int64_t s) {
    line(o, nodetag ? nodetag[s] : s + 1, 1, nodex + 3 * s, nodef + 3 * s, nodev + 3 * s);
  });
  if (std::fclose(fp) != 0) ok = false;
  return ok ? CAC_OK : io_fail(CAC_ERR_IO, "Cannot write dump file " + p);
}

extern "C" int cac_dump_element_write(const char *path, int append, int64_t step, int64_t ne, const double *nodex,
                                      const double *nodev, const double *nodef) {
  if (!path || ne < 0 || (ne && (!nodex || !nodev || !nodef)))
    return io_fail(CAC_ERR_ARG, "cac_dump_element_write: bad arguments");
  const std::string p = step_path(path, step);
  FILE *fp = std::fopen(p.c_str(), append ? "ab" : "wb");
  if (!fp) return io_fail(CAC_ERR_IO, "Cannot open dump file " + p);
  // DumpElement::header_item (dump_element.cpp:72-84)
  Out h;
  h.str("title=\"results for cac\"\n");
  h.str("variables=\"x\", \"y\", \"z\", \"vx\", \"vy\", \"vz\", \"fx\", \"fy\", \"fz\"\n");
  h.str("zone t=\"load step ");
  h.i(step);
  h.str("\", n=  ");
  h.i(8 * ne);
  h.str(" e =  ");
  h.i(ne);
  h.str(" datapacking=point, zonetype=febrick\n");
  bool ok = std::fwrite(h.s.data(), 1, h.s.size(), fp) == h.s.size();
  // node lines (epack_noscale_noimage / convert_noimage, :103-126, 156-172)
  ok = ok && write_rows(fp, 8 * ne, [&](Out &o, int64_t s) {
    for (const double *q : {nodex, nodev, nodef})
      for (int k = 0; k < 3; ++k) {
        if (q != nodex || k) o.c(' ');
        o.g6(q[3 * s + k]);
      }
    o.c('\n');
  });
  // connectivity: 1-based node lines of each element
  ok = ok && write_rows(fp, ne, [&](Out &o, int64_t e) {
    for (int k = 0; k < 8; ++k) {
      if (k) o.c(' ');
      o.i(8 * e + k + 1);
    }
    o.c('\n');
  });
  if (std::fclose(fp) != 0) ok = false;
  return ok ? CAC_OK : io_fail(CAC_ERR_IO, "Cannot write dump file " + p);
}
