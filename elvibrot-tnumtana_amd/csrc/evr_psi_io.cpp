// Wavepacket files in the reference's formats (host code).
//
// Writer: sub_save_psi (Source_ElVibRot/sub_WP/sub_module_psi_io.f90:365-424)
// with Write_header_saveFile_psi (:1441-1498), Write_list_nDindBasis
// (:1499-1604) and Write_Psi_nDBasis (:1607-1778), for a basis without
// inactive harmonic functions (nb_bi = 1) and nb_tot = nb_baie:
//   version 0, formatted (the reference's default FilePsiVersion = 0):
//     "nb_psi nb_tot", then per wavepacket one line per function
//     "n_1 .. n_D i_h i_e Re [Im]" (i_e outer, packed index inner) and
//     "end wp k";
//   version 1, formatted: namelist headerFile (Version_File, nb_psi,
//     nb_tot), the index list "n_1 .. n_D i_h i_e", then each wavepacket as
//     list-directed values (complex as "(re,im)");
//   version 1, unformatted: Fortran sequential records (4-byte length
//     markers): 'X', nb_psi, nb_tot, 1, 'X', one record of D+2 integers per
//     function, one record of nb_tot real(8) / complex(8) per wavepacket.
// n_d are the 1-based 1D function indices of the packed function
// (Rec_ndim_index), i_h = 1, i_e the electronic state.
//
// Reader: Read_header_saveFile_psi (:1294-1439), Read_psi_nDBasis
// (:920-1141) and Read_list_nDindBasis1_TO_nDindBasis2 (:1143-1292): the
// form is detected from the header; functions are matched by their index
// tuple, so a file written on another basis is projected onto this one
// (functions absent here are dropped, functions absent there stay zero).
#include <algorithm>
#include <cctype>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <fstream>
#include <iterator>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/evr_hip.h"

namespace evr {
int set_error(int code, const std::string &msg);
}

namespace {

int fail_io(int code, const std::string &m) { return evr::set_error(code, m); }

// 1-based index tuple of every packed function, from the term tables: the
// term-local index runs first-mode-fastest over the term's 1D functions and
// the map names its packed function.
bool packed_tuples(const evr_basis_desc *bd, std::vector<int> &tup)
{
  const int D = bd->D;
  tup.assign((size_t)bd->nb_packed * D, 0);
  std::vector<char> seen(bd->nb_packed, 0);
  int64_t off = 0, nseen = 0;
  std::vector<int> nb(D), n(D);
  for (int64_t t = 0; t < bd->nterms; ++t) {
    int64_t nbt = 1;
    for (int d = 0; d < D; ++d) {
      nb[d] = bd->nb[d * bd->nlev + bd->term_levels[t * D + d]];
      nbt *= nb[d];
    }
    std::fill(n.begin(), n.end(), 0);
    for (int64_t e = 0; e < nbt; ++e) {
      const int pk = bd->map[off + e];
      if (pk >= 0 && !seen[pk]) {
        seen[pk] = 1;
        ++nseen;
        for (int d = 0; d < D; ++d) tup[(size_t)pk * D + d] = n[d] + 1;
      }
      for (int d = 0; d < D; ++d) {
        if (++n[d] < nb[d]) break;
        n[d] = 0;
      }
    }
    off += nbt;
  }
  return nseen == bd->nb_packed;
}

std::string fmt_real(double v)
{
  char b[40];
  std::snprintf(b, sizeof b, "%.17E", v);
  return b;
}

// Fortran sequential unformatted record: length, payload, length.
void put_record(std::ofstream &f, const void *p, uint32_t bytes)
{
  f.write(reinterpret_cast<const char *>(&bytes), 4);
  f.write(reinterpret_cast<const char *>(p), bytes);
  f.write(reinterpret_cast<const char *>(&bytes), 4);
}

bool get_record(std::ifstream &f, std::vector<char> &rec)
{
  uint32_t n = 0, m = 0;
  if (!f.read(reinterpret_cast<char *>(&n), 4)) return false;
  rec.resize(n);
  if (n && !f.read(rec.data(), n)) return false;
  if (!f.read(reinterpret_cast<char *>(&m), 4)) return false;
  return m == n;
}

// Numbers of list-directed input: values are separated by blanks, commas or
// line ends, a slash ends the input; "(re, im)" is a complex constant (two
// numbers), "r*c" repeats the value or complex constant c r times; Fortran
// D exponents are accepted.
bool lexed_numbers(const std::string &text, std::vector<double> &out)
{
  auto sep = [](char c) { return c == ' ' || c == ',' || c == '\n' || c == '\r' || c == '\t'; };
  auto num = [](std::string t, double &v) {
    const size_t a = t.find_first_not_of(" \t\r\n"), b = t.find_last_not_of(" \t\r\n");
    if (a == std::string::npos) return false;
    t = t.substr(a, b - a + 1);
    for (char &c : t)
      if (c == 'd' || c == 'D') c = 'E';
    char *end = nullptr;
    v = std::strtod(t.c_str(), &end);
    return end != t.c_str() && *end == '\0';
  };
  const size_t n = text.size();
  size_t i = 0;
  while (i < n) {
    while (i < n && sep(text[i])) ++i;
    if (i >= n || text[i] == '/') break;
    long rep = 1;
    size_t j = i;
    while (j < n && std::isdigit((unsigned char)text[j])) ++j;
    if (j > i && j < n && text[j] == '*') {
      rep = std::strtol(text.substr(i, j - i).c_str(), nullptr, 10);
      i = j + 1;
    }
    if (rep < 1 || i >= n) return false;
    if (text[i] == '(') {
      const size_t k = text.find(')', i), c = text.find(',', i);
      if (k == std::string::npos || c == std::string::npos || c > k) return false;
      double re, im;
      if (!num(text.substr(i + 1, c - i - 1), re) || !num(text.substr(c + 1, k - c - 1), im)) return false;
      for (long r = 0; r < rep; ++r) {
        out.push_back(re);
        out.push_back(im);
      }
      i = k + 1;
    } else {
      size_t k = i;
      while (k < n && !sep(text[k]) && text[k] != '/') ++k;
      double v;
      if (!num(text.substr(i, k - i), v)) return false;
      for (long r = 0; r < rep; ++r) out.push_back(v);
      i = k;
    }
  }
  return true;
}

std::string lower(std::string s)
{
  for (char &c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

}  // namespace

namespace {

int psi_save(const evr_basis_desc *bd, int nb0, int nwp, int cplx, const double *psi, int version,
             int formatted, const char *path)
{
  if (!bd || !psi || !path || nb0 < 1 || nwp < 1) return fail_io(EVR_ERR_ARG, "bad argument");
  if (version != 0 && version != 1) return fail_io(EVR_ERR_ARG, "version must be 0 or 1");
  if (version == 0 && !formatted)
    return fail_io(EVR_ERR_ARG, "an unformatted file is written in version 1 (as the reference)");
  const int D = bd->D;
  std::vector<int> tup;
  if (!packed_tuples(bd, tup)) return fail_io(EVR_ERR_ARG, "map does not cover the packed basis");
  const int64_t nb = bd->nb_packed, ntot = nb * nb0;
  const int ncol = cplx ? 2 : 1;
  auto val = [&](int w, int64_t row, int part) { return psi[(size_t)(w * ncol + part) * ntot + row]; };
  if (formatted) {
    std::ofstream f(path);
    if (!f) return fail_io(EVR_ERR_IO, std::string("cannot open ") + path);
    auto index_line = [&](int64_t row) {
      const int64_t pk = row % nb;
      const int ie = (int)(row / nb) + 1;
      std::string s;
      for (int d = 0; d < D; ++d) s += " " + std::to_string(tup[(size_t)pk * D + d]);
      s += " 1 " + std::to_string(ie);
      return s;
    };
    if (version == 0) {
      f << " " << nwp << " " << ntot << "\n";
      for (int w = 0; w < nwp; ++w) {
        for (int64_t r = 0; r < ntot; ++r) {
          f << index_line(r) << " " << fmt_real(val(w, r, 0));
          if (cplx) f << " " << fmt_real(val(w, r, 1));
          f << "\n";
        }
        f << " end wp " << (w + 1) << "\n";
      }
    } else {
      f << " &HEADERFILE\n VERSION_FILE=1,\n NB_PSI=" << nwp << ",\n NB_TOT=" << ntot << ",\n /\n";
      for (int64_t r = 0; r < ntot; ++r) f << index_line(r) << "\n";
      for (int w = 0; w < nwp; ++w)
        for (int64_t r = 0; r < ntot; ++r) {
          if (cplx)
            f << " (" << fmt_real(val(w, r, 0)) << "," << fmt_real(val(w, r, 1)) << ")\n";
          else
            f << " " << fmt_real(val(w, r, 0)) << "\n";
        }
    }
    if (!f) return fail_io(EVR_ERR_IO, std::string("write failed: ") + path);
    return EVR_OK;
  }
  // default 4-byte integers and record markers, as a reference build writes them
  if (ntot > INT32_MAX || 8.0 * ntot * ncol > 2147483647.0)
    return fail_io(EVR_ERR_SIZE, "wavepacket too large for an unformatted file with 4-byte markers");
  std::ofstream f(path, std::ios::binary);
  if (!f) return fail_io(EVR_ERR_IO, std::string("cannot open ") + path);
  const char X = 'X';
  const int32_t hdr[3] = {nwp, (int32_t)ntot, 1};
  put_record(f, &X, 1);
  for (int k = 0; k < 3; ++k) put_record(f, &hdr[k], 4);
  put_record(f, &X, 1);
  std::vector<int32_t> ind(D + 2);
  for (int64_t r = 0; r < ntot; ++r) {
    const int64_t pk = r % nb;
    for (int d = 0; d < D; ++d) ind[d] = tup[(size_t)pk * D + d];
    ind[D] = 1;
    ind[D + 1] = (int32_t)(r / nb) + 1;
    put_record(f, ind.data(), (uint32_t)(4 * (D + 2)));
  }
  std::vector<double> buf((size_t)ntot * ncol);
  for (int w = 0; w < nwp; ++w) {
    for (int64_t r = 0; r < ntot; ++r)
      for (int c = 0; c < ncol; ++c) buf[(size_t)r * ncol + c] = val(w, r, c);
    put_record(f, buf.data(), (uint32_t)(8 * buf.size()));
  }
  if (!f) return fail_io(EVR_ERR_IO, std::string("write failed: ") + path);
  return EVR_OK;
}

int psi_read(const evr_basis_desc *bd, int nb0, int cplx, int max_wp, double *psi, int *nwp,
             const char *path)
{
  if (!bd || !psi || !path || !nwp || nb0 < 1 || max_wp < 1) return fail_io(EVR_ERR_ARG, "bad argument");
  *nwp = 0;
  const int D = bd->D;
  std::vector<int> tup;
  if (!packed_tuples(bd, tup)) return fail_io(EVR_ERR_ARG, "map does not cover the packed basis");
  const int64_t nb = bd->nb_packed, ntot = nb * nb0;
  const int ncol = cplx ? 2 : 1;
  std::fill(psi, psi + (size_t)ntot * ncol * max_wp, 0.0);
  std::map<std::vector<int>, int64_t> row_of;  // (n_1..n_D, i_h, i_e) -> row of this basis
  for (int64_t r = 0; r < ntot; ++r) {
    std::vector<int> k(tup.begin() + (r % nb) * D, tup.begin() + (r % nb + 1) * D);
    k.push_back(1);
    k.push_back((int)(r / nb) + 1);
    row_of.emplace(std::move(k), r);
  }
  auto row_for = [&](const std::vector<int> &key) -> int64_t {
    auto it = row_of.find(key);
    return it == row_of.end() ? -1 : it->second;
  };
  auto put = [&](int w, int64_t row, double re, double im) {
    if (row < 0 || w >= max_wp) return;
    psi[(size_t)(w * ncol) * ntot + row] = re;
    if (cplx) psi[(size_t)(w * ncol + 1) * ntot + row] = im;
  };

  // unformatted: the first record is the 1-byte 'X'
  {
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail_io(EVR_ERR_IO, std::string("cannot open ") + path);
    uint32_t n0 = 0;
    char c = 0;
    if (f.read(reinterpret_cast<char *>(&n0), 4) && n0 == 1 && f.read(&c, 1) && c == 'X') {
      f.seekg(0);
      std::vector<char> rec;
      int32_t h[3];
      if (!get_record(f, rec)) return fail_io(EVR_ERR_IO, "bad record");
      for (int k = 0; k < 3; ++k) {
        if (!get_record(f, rec) || rec.size() != 4) return fail_io(EVR_ERR_IO, "bad header record");
        std::memcpy(&h[k], rec.data(), 4);
      }
      if (!get_record(f, rec)) return fail_io(EVR_ERR_IO, "bad header record");
      if (h[2] != 1) return fail_io(EVR_ERR_IO, "unformatted wavepacket files are version 1");
      const int npsi = h[0];
      const int64_t ftot = h[1];
      // counts from the file are checked before anything is sized from them:
      // every index record takes 4 (D + 2) + 8 bytes of the file
      const std::streampos here = f.tellg();
      f.seekg(0, std::ios::end);
      const int64_t fsize = (int64_t)f.tellg();
      f.seekg(here);
      if (npsi < 0 || ftot < 0 || ftot > fsize / (4 * (D + 2) + 8))
        return fail_io(EVR_ERR_IO, "header counts (nb_psi, nb_tot) do not fit the file");
      std::vector<int64_t> rows(ftot);
      std::vector<int> key(D + 2);
      for (int64_t i = 0; i < ftot; ++i) {
        if (!get_record(f, rec) || rec.size() != 4 * (size_t)(D + 2))
          return fail_io(EVR_ERR_IO, "index record does not hold D + 2 integers");
        for (int d = 0; d < D + 2; ++d) std::memcpy(&key[d], rec.data() + 4 * d, 4);
        rows[i] = row_for(key);
      }
      for (int w = 0; w < npsi; ++w) {
        if (!get_record(f, rec) || rec.size() != 8 * (size_t)ftot * ncol)
          return fail_io(EVR_ERR_IO, "wavepacket record of the wrong size (real/complex?)");
        const double *v = reinterpret_cast<const double *>(rec.data());
        for (int64_t i = 0; i < ftot; ++i) put(w, rows[i], v[i * ncol], cplx ? v[i * ncol + 1] : 0.0);
      }
      *nwp = npsi < max_wp ? npsi : max_wp;
      return EVR_OK;
    }
  }

  std::ifstream f(path);
  std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  std::istringstream in(text);
  std::string line;
  if (!std::getline(in, line)) return fail_io(EVR_ERR_IO, "empty file");
  const std::string l0 = lower(line);
  if (l0.find("&headerfile") != std::string::npos) {
    // namelist header up to '/': version, nb_psi, nb_tot
    std::string nml = l0;
    while (nml.find('/') == std::string::npos && std::getline(in, line)) nml += " " + lower(line);
    nml = nml.substr(nml.find("&headerfile") + 11);
    nml = nml.substr(0, nml.find('/'));
    for (char &c : nml)
      if (c == ',' || c == '=') c = ' ';
    std::istringstream ns(nml);
    std::string k;
    long version = 0, npsi = 0, ftot = 0;
    while (ns >> k) {
      long v = 0;
      if (!(ns >> v)) return fail_io(EVR_ERR_IO, "bad headerFile namelist");
      if (k == "version_file") version = v;
      else if (k == "nb_psi") npsi = v;
      else if (k == "nb_tot") ftot = v;
    }
    if (version != 1) return fail_io(EVR_ERR_IO, "namelist header with Version_File /= 1");
    // every index line holds D + 2 integers (at least 2 characters each)
    if (npsi < 0 || ftot < 0 || ftot > (long)(text.size() / (2 * (D + 2)) + 1))
      return fail_io(EVR_ERR_IO, "header counts (nb_psi, nb_tot) do not fit the file");
    std::vector<int64_t> rows(ftot);
    std::vector<int> key(D + 2);
    for (long i = 0; i < ftot; ++i) {
      if (!std::getline(in, line)) return fail_io(EVR_ERR_IO, "index list too short");
      std::istringstream ls(line);
      for (int d = 0; d < D + 2; ++d)
        if (!(ls >> key[d])) return fail_io(EVR_ERR_IO, "index line does not hold D + 2 integers");
      rows[i] = row_for(key);
    }
    std::string rest((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    std::vector<double> v;
    if (!lexed_numbers(rest, v)) return fail_io(EVR_ERR_IO, "bad value in the wavepacket list");
    const size_t per = (size_t)ftot * ncol;
    if (v.size() < per * npsi) return fail_io(EVR_ERR_IO, "fewer values than nb_psi x nb_tot");
    for (long w = 0; w < npsi; ++w)
      for (long i = 0; i < ftot; ++i)
        put((int)w, rows[i], v[w * per + i * ncol], cplx ? v[w * per + i * ncol + 1] : 0.0);
    *nwp = npsi < max_wp ? (int)npsi : max_wp;
    return EVR_OK;
  }
  // version 0: "nb_psi [nb_tot]", then index + value lines closed by "end wp"
  long npsi = 0;
  {
    std::istringstream hs(line);
    if (!(hs >> npsi) || npsi < 1) return fail_io(EVR_ERR_IO, "header is neither a namelist nor nb_psi");
  }
  int w = 0;
  std::vector<int> key(D + 2);
  while (w < npsi && std::getline(in, line)) {
    if (lower(line).find("end") != std::string::npos) {
      ++w;
      continue;
    }
    std::vector<double> v;
    if (!lexed_numbers(line, v)) return fail_io(EVR_ERR_IO, "bad wavepacket line: " + line);
    if (v.empty()) continue;
    if (v.size() != (size_t)(D + 2 + ncol))
      return fail_io(EVR_ERR_IO, "wavepacket line does not hold D + 2 indices and the value(s)");
    for (int d = 0; d < D + 2; ++d) key[d] = (int)v[d];
    put(w, row_for(key), v[D + 2], cplx ? v[D + 3] : 0.0);
  }
  *nwp = (int)std::min<long>(w, max_wp);
  return EVR_OK;
}

}  // namespace

extern "C" {

// The entry points return a status for any failure, including an allocation
// that a malformed file or an oversized basis makes fail.
int evr_psi_save(const evr_basis_desc *bd, int nb0, int nwp, int cplx, const double *psi, int version,
                 int formatted, const char *path)
{
  try {
    return psi_save(bd, nb0, nwp, cplx, psi, version, formatted, path);
  } catch (const std::exception &e) {
    return fail_io(EVR_ERR_IO, std::string("evr_psi_save: ") + e.what());
  }
}

int evr_propa_save_restart(double T, int nwp, int64_t size_wp, const double *psi, const char *path)
{
  if (!psi || !path || nwp < 1 || size_wp < 1) return fail_io(EVR_ERR_ARG, "bad argument");
  try {
    std::ofstream f(path);
    if (!f) return fail_io(EVR_ERR_IO, std::string("cannot open ") + path);
    // write(no_restart,*) T, size(WP), size(WP(1)%CvecB), then each CvecB
    f << " " << fmt_real(T) << " " << nwp << " " << size_wp << "\n";
    for (int w = 0; w < nwp; ++w) {
      const double *re = psi + (size_t)(2 * w) * size_wp, *im = psi + (size_t)(2 * w + 1) * size_wp;
      for (int64_t i = 0; i < size_wp; ++i) f << " (" << fmt_real(re[i]) << "," << fmt_real(im[i]) << ")\n";
    }
    if (!f) return fail_io(EVR_ERR_IO, std::string("write failed: ") + path);
    return EVR_OK;
  } catch (const std::exception &e) {
    return fail_io(EVR_ERR_IO, std::string("evr_propa_save_restart: ") + e.what());
  }
}

int evr_propa_read_restart(const char *path, int nwp, int64_t size_wp, double *psi, double *T, int *found)
{
  if (!path || !psi || !T || !found || nwp < 1 || size_wp < 1) return fail_io(EVR_ERR_ARG, "bad argument");
  try {
    *T = 0.0;
    *found = 0;
    std::ifstream f(path);
    if (!f) return fail_io(EVR_ERR_IO, std::string("cannot open ") + path);
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<double> v;
    // read(no_restart,*,IOSTAT=err_read) T, nb_WP_file, size_WP_file: a file
    // whose first record does not hold them means no restart (T0 = 0)
    const size_t eol = text.find('\n');
    if (!lexed_numbers(text.substr(0, eol), v) || v.size() < 3) return EVR_OK;
    const double t0 = v[0];
    const double nwf = v[1], szf = v[2];
    if (nwf != (double)nwp || szf != (double)size_wp)
      return fail_io(EVR_ERR_IO, "restart file holds " + std::to_string((long long)nwf) + " wavepackets of " +
                                     std::to_string((long long)szf) + " values, not " + std::to_string(nwp) +
                                     " of " + std::to_string((long long)size_wp));
    v.clear();
    if (eol == std::string::npos || !lexed_numbers(text.substr(eol + 1), v))
      return fail_io(EVR_ERR_IO, "bad wavepacket values in the restart file");
    if (v.size() < (size_t)(2 * nwp) * size_wp) return fail_io(EVR_ERR_IO, "restart file too short");
    for (int w = 0; w < nwp; ++w)
      for (int64_t i = 0; i < size_wp; ++i) {
        psi[(size_t)(2 * w) * size_wp + i] = v[2 * ((size_t)w * size_wp + i)];
        psi[(size_t)(2 * w + 1) * size_wp + i] = v[2 * ((size_t)w * size_wp + i) + 1];
      }
    *T = t0;
    *found = 1;
    return EVR_OK;
  } catch (const std::exception &e) {
    return fail_io(EVR_ERR_IO, std::string("evr_propa_read_restart: ") + e.what());
  }
}

int evr_psi_read(const evr_basis_desc *bd, int nb0, int cplx, int max_wp, double *psi, int *nwp,
                 const char *path)
{
  try {
    return psi_read(bd, nb0, cplx, max_wp, psi, nwp, path);
  } catch (const std::exception &e) {
    if (nwp) *nwp = 0;
    return fail_io(EVR_ERR_IO, std::string("evr_psi_read: ") + e.what());
  }
}

}  // extern "C"
