// FITS map I/O (s2let_fits.c:9-551): MW / MWSS maps and spin (Q, U) maps as one
// binary-table extension of double columns after an empty primary HDU, with the
// keywords L (and MWSS = 1 for MWSS maps); readers take the second HDU's first (and
// second) column, converting 4- or 8-byte reals, and NSIDE of HEALPix files.
// Self-contained (FITS standard 4.0: 2880-byte records of 80-character cards,
// big-endian data), so the library needs no FITS library at run time.
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace {

constexpr size_t kRecord = 2880;
// status codes of the reference's FITS library, used as its exit codes on read errors
constexpr int kFileNotOpened = 104, kKeyNoExist = 202, kBadFormat = 261;

[[noreturn]] void fail(const char *fn, const std::string &what, int status) {
  fprintf(stderr, "s2let FITS: %s: %s (status %d)\n", fn, what.c_str(), status);
  exit(status);
}

std::string card(const std::string &key, const std::string &value, const std::string &comment) {
  char b[81];
  snprintf(b, sizeof b, "%-8.8s= %s / %s", key.c_str(), value.c_str(), comment.c_str());
  std::string s(b);
  s.resize(80, ' ');
  return s;
}
std::string num(long v) {
  char b[32];
  snprintf(b, sizeof b, "%20ld", v);
  return b;
}
std::string str(const std::string &v) {
  std::string q = "'" + v;
  while (q.size() < 9) q += ' ';
  q += "'";
  while (q.size() < 20) q += ' ';
  return q;
}
std::string bare(const std::string &text) {
  std::string s = text;
  s.resize(80, ' ');
  return s;
}
void pad(std::string &h, char c) {
  while (h.size() % kRecord) h += c;
}

void put_be64(unsigned char *p, double v) {
  uint64_t u;
  memcpy(&u, &v, 8);
  for (int i = 7; i >= 0; --i, u >>= 8) p[i] = (unsigned char)(u & 0xff);
}
double get_be64(const unsigned char *p) {
  uint64_t u = 0;
  for (int i = 0; i < 8; ++i) u = (u << 8) | p[i];
  double v;
  memcpy(&v, &u, 8);
  return v;
}
float get_be32(const unsigned char *p) {
  uint32_t u = 0;
  for (int i = 0; i < 4; ++i) u = (u << 8) | p[i];
  float v;
  memcpy(&v, &u, 4);
  return v;
}

// empty primary HDU + a binary table of ncol double columns named names[], npix rows
void write_table(const char *fn, const char *filename, const std::vector<const double *> &cols,
                 const std::vector<std::string> &names, long npix, long L, bool mwss) {
  std::string h;
  h += card("SIMPLE", num(1).replace(19, 1, "T"), "file does conform to FITS standard");
  h += card("BITPIX", num(16), "number of bits per data pixel");
  h += card("NAXIS", num(0), "number of data axes");
  h += card("EXTEND", num(1).replace(19, 1, "T"), "FITS dataset may contain extensions");
  h += bare("END");
  pad(h, ' ');
  const int nc = (int)cols.size();
  h += card("XTENSION", str("BINTABLE"), "binary table extension");
  h += card("BITPIX", num(8), "8-bit bytes");
  h += card("NAXIS", num(2), "2-dimensional binary table");
  h += card("NAXIS1", num(8L * nc), "width of table in bytes");
  h += card("NAXIS2", num(npix), "number of rows in table");
  h += card("PCOUNT", num(0), "size of special data area");
  h += card("GCOUNT", num(1), "one data group (required keyword)");
  h += card("TFIELDS", num(nc), "number of fields in each row");
  for (int c = 0; c < nc; ++c) {
    const std::string k = std::to_string(c + 1);
    h += card("TTYPE" + k, str(names[c]), "label for field " + k);
    h += card("TFORM" + k, str("1D"), "data format of field: 8-byte DOUBLE");
    h += card("TUNIT" + k, str(" "), "physical unit of field");
  }
  h += card("EXTNAME", str("BINTABLE"), "name of this binary table extension");
  h += card("L", num(L), "Resolution parameter");
  if (mwss) h += card("MWSS", num(1), "Sampling scheme parameter");
  h += bare("END");
  pad(h, ' ');
  std::vector<unsigned char> data((size_t)npix * nc * 8);
  for (long r = 0; r < npix; ++r)
    for (int c = 0; c < nc; ++c) put_be64(&data[((size_t)r * nc + c) * 8], cols[c][r]);
  while (data.size() % kRecord) data.push_back(0);
  FILE *fp = fopen(filename, "wb");
  if (!fp) {
    fprintf(stderr, "%s: could not create new fits file %s.\n", fn, filename);
    return;
  }
  fwrite(h.data(), 1, h.size(), fp);
  fwrite(data.data(), 1, data.size(), fp);
  fclose(fp);
}

struct Hdu {
  std::map<std::string, std::string> keys;
  size_t data_off = 0;
};

std::string trim(const std::string &s) {
  size_t a = s.find_first_not_of(' '), b = s.find_last_not_of(' ');
  return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

// the first two HDUs of a FITS file: keywords and data offsets
bool read_hdus(const std::vector<unsigned char> &f, Hdu hdu[2]) {
  size_t off = 0;
  for (int k = 0; k < 2; ++k) {
    bool end = false;
    while (!end) {
      if (off + kRecord > f.size()) return false;
      for (size_t c = 0; c < kRecord; c += 80) {
        const std::string cd((const char *)&f[off + c], 80);
        const std::string key = trim(cd.substr(0, 8));
        if (key == "END") {
          end = true;
          break;
        }
        if (cd.size() > 9 && cd[8] == '=') {
          std::string v = cd.substr(10);
          if (trim(v).size() && trim(v)[0] == '\'') {
            const size_t a = v.find('\''), b = v.find('\'', a + 1);
            v = trim(v.substr(a + 1, b == std::string::npos ? std::string::npos : b - a - 1));
          } else {
            const size_t sl = v.find('/');
            v = trim(v.substr(0, sl));
          }
          hdu[k].keys[key] = v;
        }
      }
      off += kRecord;
    }
    hdu[k].data_off = off;
    // data size: |BITPIX|/8 * GCOUNT * (PCOUNT + prod NAXISi)
    auto geti = [&](const char *key, long def) {
      auto it = hdu[k].keys.find(key);
      return it == hdu[k].keys.end() ? def : atol(it->second.c_str());
    };
    const long naxis = geti("NAXIS", 0);
    long n = naxis > 0 ? 1 : 0;
    for (long i = 1; i <= naxis; ++i) n *= geti(("NAXIS" + std::to_string(i)).c_str(), 0);
    const size_t bytes = (size_t)(labs(geti("BITPIX", 8)) / 8) * geti("GCOUNT", 1) * (geti("PCOUNT", 0) + n);
    off += (bytes + kRecord - 1) / kRecord * kRecord;
  }
  return true;
}

std::vector<unsigned char> slurp(const char *fn, const char *filename) {
  FILE *fp = fopen(filename, "rb");
  if (!fp) fail(fn, std::string("could not open ") + filename, kFileNotOpened);
  std::vector<unsigned char> f;
  unsigned char buf[1 << 16];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, fp)) > 0) f.insert(f.end(), buf, buf + n);
  fclose(fp);
  return f;
}

long key_long(const char *fn, const Hdu &h, const char *key) {
  auto it = h.keys.find(key);
  if (it == h.keys.end()) fail(fn, std::string("keyword ") + key + " not found", kKeyNoExist);
  return atol(it->second.c_str());
}

// reads npix values of column col (1-based) of the binary table (NAXIS2 rows of
// npix / NAXIS2 elements each, 4- or 8-byte reals)
void read_column(const char *fn, const std::vector<unsigned char> &f, const Hdu &h, int col, double *out, long npix) {
  const long width = key_long(fn, h, "NAXIS1"), rows = key_long(fn, h, "NAXIS2");
  const long nf = key_long(fn, h, "TFIELDS");
  if (rows <= 0 || npix % rows != 0) fprintf(stderr, "%s: Problem with npix.\n", fn);
  const long per = rows > 0 ? npix / rows : 0;
  long cofs = 0;
  int esz = 8;
  bool real8 = true;
  for (long c = 1; c <= nf; ++c) {
    const std::string t = h.keys.count("TFORM" + std::to_string(c)) ? h.keys.at("TFORM" + std::to_string(c)) : "";
    const long rep = (t.size() > 1 && isdigit((unsigned char)t[0])) ? atol(t.c_str()) : 1;
    const char code = t.empty() ? 'D' : t[t.size() - 1];
    const int sz = code == 'D' ? 8 : code == 'E' ? 4 : code == 'K' ? 8 : code == 'J' ? 4 : code == 'I' ? 2 : 1;
    if (c == col) {
      if (code != 'D' && code != 'E') fail(fn, "column is not real", kBadFormat);
      esz = sz;
      real8 = code == 'D';
      if (rep < per) fail(fn, "column narrower than the map", kBadFormat);
      break;
    }
    cofs += rep * sz;
  }
  const size_t need = h.data_off + (size_t)rows * width;
  if (need > f.size()) fail(fn, "file shorter than its table", kBadFormat);
  for (long r = 0; r < rows; ++r)
    for (long e = 0; e < per; ++e) {
      const unsigned char *p = &f[h.data_off + (size_t)r * width + cofs + (size_t)e * esz];
      out[r * per + e] = real8 ? get_be64(p) : (double)get_be32(p);
    }
}

void read_maps(const char *fn, const char *filename, int L, bool mwss, double *a, double *b, bool spin_msgs) {
  const std::vector<unsigned char> f = slurp(fn, filename);
  Hdu h[2];
  if (!read_hdus(f, h)) fail(fn, "truncated FITS file", kBadFormat);
  if (!h[1].keys.count("XTENSION") || h[1].keys.at("XTENSION") != "BINTABLE")
    fprintf(stderr, "%s: Extension is not binary!\n", fn);
  const long Lread = key_long(fn, h[1], "L");
  if (Lread != L) printf("Attention : read L = %li but you specified L = %i\n", Lread, L);
  if (mwss) {
    const long m = key_long(fn, h[1], "MWSS");
    if (m != 1) printf("Attention : the file is not stored with MWSS sampling.\n");
  }
  const long npix = mwss ? (long)(L + 1) * 2 * L : (long)L * (2 * L - 1);
  if (spin_msgs) {
    printf("naxis[0] = %li\n", key_long(fn, h[1], "NAXIS1"));
    printf("naxis[1] = %li\n", key_long(fn, h[1], "NAXIS2"));
  }
  read_column(fn, f, h[1], 1, a, npix);
  if (b) read_column(fn, f, h[1], 2, b, npix);
}

}  // namespace

extern "C" {

int s2let_fits_mw_read_bandlimit(char *filename) {
  const std::vector<unsigned char> f = slurp(__func__, filename);
  Hdu h[2];
  if (!read_hdus(f, h)) fail(__func__, "truncated FITS file", kBadFormat);
  if (!h[1].keys.count("XTENSION") || h[1].keys.at("XTENSION") != "BINTABLE")
    fprintf(stderr, "%s: Extension is not binary!\n", __func__);
  auto it = h[1].keys.find("L");
  return it == h[1].keys.end() ? -1 : atoi(it->second.c_str());
}

int s2let_fits_hpx_read_nside(char *filename) {
  const std::vector<unsigned char> f = slurp(__func__, filename);
  Hdu h[2];
  if (!read_hdus(f, h)) fail(__func__, "truncated FITS file", kBadFormat);
  if (!h[1].keys.count("XTENSION") || h[1].keys.at("XTENSION") != "BINTABLE")
    fprintf(stderr, "%s: Extension is not binary!\n", __func__);
  auto it = h[1].keys.find("NSIDE");
  return it == h[1].keys.end() ? -1 : atoi(it->second.c_str());
}

void s2let_fits_mw_write_map(char *filename, double *f, int L) {
  write_table(__func__, filename, {f}, {"SIGNAL"}, (long)L * (2 * L - 1), L, false);
}
void s2let_fits_mwss_write_map(char *filename, double *f, int L) {
  write_table(__func__, filename, {f}, {"SIGNAL"}, (long)(L + 1) * 2 * L, L, true);
}
void s2let_fits_mw_read_map(double *f, char *filename, int L) { read_maps(__func__, filename, L, false, f, nullptr, false); }
void s2let_fits_mwss_read_map(double *f, char *filename, int L) {
  read_maps(__func__, filename, L, true, f, nullptr, false);
}
void s2let_fits_mw_write_spin_maps(char *filename, double *fQ, double *fU, int L) {
  write_table(__func__, filename, {fQ, fU}, {"Q_POLARISATION", "U_POLARISATION"}, (long)L * (2 * L - 1), L, false);
}
void s2let_fits_mwss_write_spin_maps(char *filename, double *fQ, double *fU, int L) {
  write_table(__func__, filename, {fQ, fU}, {"Q_POLARISATION", "U_POLARISATION"}, (long)(L + 1) * 2 * L, L, true);
}
void s2let_fits_mw_read_spin_maps(double *fQ, double *fU, char *filename, int L) {
  read_maps(__func__, filename, L, false, fQ, fU, true);
}
void s2let_fits_mwss_read_spin_maps(double *fQ, double *fU, char *filename, int L) {
  read_maps(__func__, filename, L, true, fQ, fU, true);
}

}  // extern "C"
