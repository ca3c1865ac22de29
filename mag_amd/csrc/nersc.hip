// NERSC gauge-configuration I/O: export_config_nersc / import_config_nersc
// (field-matrix.h:445-713). Host code; the device is used only through the
// library's own entry points (download/upload, plaquette, reunitarize).
//
// File layout (as the reference writes it):
//   BEGIN_HEADER
//   HDR_VERSION = 1.0
//   DATATYPE = 4D_SU3_GAUGE | 4D_SU3_GAUGE_3x3
//   DIMENSION_1..4, CRC32HASH, CHECKSUM (hex), LINK_TRACE, PLAQUETTE,
//   CREATOR, ARCHIVE_DATE, ENSEMBLE_LABEL, FLOATING_POINT = IEEE64BIG,
//   ENSEMBLE_ID, SEQUENCE_NUMBER, BETA
//   END_HEADER
// then for every site in lexicographic order (x0 fastest) and mu = 0..3 the
// link row-major as (re, im) IEEE64 big-endian: 3 rows (18 reals) or the
// first 2 (12 reals, 4D_SU3_GAUGE). Floating header values are printed as a
// C++ ostream prints them: 12 significant digits in the two-row format (the
// reference sets precision(12) there, field-matrix.h:490), 6 otherwise.
// CHECKSUM is the sum mod 2^32 of the data taken as 32-bit words of the
// host (little-endian) doubles, i.e. of the high and low halves of every
// double (the order of the halves does not change the sum). That is what the
// reference's import checks for both DATATYPEs (field-matrix.h:684, 697) and
// what its export writes for 4D_SU3_GAUGE_3x3 (field-matrix.h:527-529, before
// the byte swap). For 4D_SU3_GAUGE its export sums after the byte swap
// (field-matrix.h:473-477), i.e. the stored big-endian bytes taken as
// little-endian words, so its own import refuses those files. The writer here
// always writes the host-order sum (readable by the reference's import); the
// reader accepts either sum for 4D_SU3_GAUGE and reports which one matched.
// CRC32HASH here is the IEEE CRC-32 (zlib's crc32) of the data in host byte
// order, the bytes the reference hands to qlat's field_hash_crc32 before its
// byte swap (field-matrix.h:470, 523); field_hash_crc32 itself is not part of
// the reference sources and the reference's import does not check it.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "runtime.hpp"

using namespace mag;

namespace {

uint32_t crc32_ieee(const unsigned char* p, size_t n) {
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t c = i;
      for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      table[i] = c;
    }
    init = true;
  }
  uint32_t c = 0xFFFFFFFFu;
  for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
  return c ^ 0xFFFFFFFFu;
}

// Host doubles <-> big-endian bytes (the host is little-endian x86-64).
void swap64(unsigned char* p, size_t ndouble) {
  for (size_t i = 0; i < ndouble; ++i, p += 8)
    for (int k = 0; k < 4; ++k) std::swap(p[k], p[7 - k]);
}

uint32_t checksum32(const double* d, size_t n) {
  uint32_t s = 0;
  for (size_t i = 0; i < n; ++i) {
    uint64_t w;
    std::memcpy(&w, d + i, 8);
    s += (uint32_t)(w & 0xFFFFFFFFu) + (uint32_t)(w >> 32);
  }
  return s;
}

// The same sum over the byte-swapped doubles (the stored big-endian bytes
// taken as little-endian 32-bit words): each half byte-reversed.
uint32_t checksum32_swapped(const double* d, size_t n) {
  uint32_t s = 0;
  for (size_t i = 0; i < n; ++i) {
    uint64_t w;
    std::memcpy(&w, d + i, 8);
    s += __builtin_bswap32((uint32_t)(w & 0xFFFFFFFFu)) + __builtin_bswap32((uint32_t)(w >> 32));
  }
  return s;
}

std::string gfmt(double v, int prec) {
  char b[64];
  std::snprintf(b, sizeof b, "%.*g", prec, v);
  return b;
}

// "KEY = value" lines; returns false when the key is absent.
bool keyword(const std::string& line, const char* key, std::string* val) {
  const size_t k = line.find(key);
  if (k == std::string::npos) return false;
  const size_t eq = line.find('=', k);
  if (eq == std::string::npos) return false;
  size_t a = eq + 1;
  while (a < line.size() && (line[a] == ' ' || line[a] == '\t')) ++a;
  size_t b = line.size();
  while (b > a && (line[b - 1] == '\n' || line[b - 1] == '\r')) --b;
  *val = line.substr(a, b - a);
  return true;
}

long volume(const int d[4]) { return (long)d[0] * d[1] * d[2] * d[3]; }

}  // namespace

extern "C" {

int mag_nersc_write(const char* path, const double* aos, mag_nersc_header* h) {
  if (!path || !aos || !h) return MAG_ERR_ARG;
  for (int i = 0; i < 4; ++i)
    if (h->dims[i] < 1) return MAG_ERR_ARG;
  const long nlink = 4 * volume(h->dims);
  const int per = h->skip_third ? 12 : 18;
  std::vector<double> data((size_t)nlink * per);
  for (long l = 0; l < nlink; ++l) std::memcpy(&data[(size_t)l * per], aos + (size_t)l * 18, per * sizeof(double));
  h->checksum = checksum32(data.data(), data.size());
  h->checksum_swapped = 0;
  h->crc32 = crc32_ieee((const unsigned char*)data.data(), data.size() * 8);
  swap64((unsigned char*)data.data(), data.size());
  const int prec = h->skip_third ? 12 : 6;
  char hex[32];
  std::string o = "BEGIN_HEADER\nHDR_VERSION = 1.0\n";
  o += h->skip_third ? "DATATYPE = 4D_SU3_GAUGE\n" : "DATATYPE = 4D_SU3_GAUGE_3x3\n";
  for (int i = 0; i < 4; ++i) o += "DIMENSION_" + std::to_string(i + 1) + " = " + std::to_string(h->dims[i]) + "\n";
  std::snprintf(hex, sizeof hex, "%x", h->crc32);
  o += std::string("CRC32HASH = ") + hex + "\n";
  std::snprintf(hex, sizeof hex, "%x", h->checksum);
  o += std::string("CHECKSUM = ") + hex + "\n";
  o += "LINK_TRACE = " + gfmt(h->link_trace, prec) + "\n";
  o += "PLAQUETTE = " + gfmt(h->plaquette, prec) + "\n";
  o += "CREATOR = RBC\n";
  const time_t now = std::time(nullptr);
  o += std::string("ARCHIVE_DATE = ") + std::ctime(&now);
  o += std::string("ENSEMBLE_LABEL = ") + h->ensemble_label + "\n";
  o += "FLOATING_POINT = IEEE64BIG\nENSEMBLE_ID = NOT yet implemented\n";
  o += "SEQUENCE_NUMBER = " + std::to_string(h->sequence_number) + "\n";
  o += "BETA = " + gfmt(h->beta, prec) + "\n";
  o += "END_HEADER\n";
  FILE* f = std::fopen(path, "wb");
  if (!f) return MAG_ERR_IO;
  bool ok = std::fwrite(o.data(), 1, o.size(), f) == o.size() &&
            std::fwrite(data.data(), 8, data.size(), f) == data.size();
  ok = (std::fclose(f) == 0) && ok;
  return ok ? MAG_OK : MAG_ERR_IO;
}

int mag_nersc_read(const char* path, double* aos, mag_nersc_header* h) {
  if (!path || !h) return MAG_ERR_ARG;
  FILE* f = std::fopen(path, "rb");
  if (!f) return MAG_ERR_IO;
  std::memset(h, 0, sizeof *h);
  std::string type, v;
  bool have_ck = false, end = false;
  char buf[4096];
  while (std::fgets(buf, sizeof buf, f)) {
    const std::string line = buf;
    if (keyword(line, "DATATYPE", &v)) type = v;
    for (int i = 0; i < 4; ++i) {
      const std::string key = "DIMENSION_" + std::to_string(i + 1);
      if (keyword(line, key.c_str(), &v)) h->dims[i] = std::atoi(v.c_str());
    }
    if (keyword(line, "CRC32HASH", &v)) h->crc32 = (uint32_t)std::strtoul(v.c_str(), nullptr, 16);
    else if (keyword(line, "CHECKSUM", &v)) {
      h->checksum = (uint32_t)std::strtoul(v.c_str(), nullptr, 16);
      have_ck = true;
    }
    if (keyword(line, "LINK_TRACE", &v)) h->link_trace = std::strtod(v.c_str(), nullptr);
    if (keyword(line, "PLAQUETTE", &v)) h->plaquette = std::strtod(v.c_str(), nullptr);
    if (keyword(line, "SEQUENCE_NUMBER", &v)) h->sequence_number = std::atoi(v.c_str());
    if (keyword(line, "ENSEMBLE_LABEL", &v))
      std::snprintf(h->ensemble_label, sizeof h->ensemble_label, "%s", v.c_str());
    if (line.compare(0, 4, "BETA") == 0 && keyword(line, "BETA", &v)) h->beta = std::strtod(v.c_str(), nullptr);
    if (line.find("END_HEADER") != std::string::npos) {
      end = true;
      break;
    }
  }
  int r = MAG_OK;
  if (!end || !have_ck || volume(h->dims) <= 0) r = MAG_ERR_IO;
  else if (type == "4D_SU3_GAUGE") h->skip_third = 1;
  else if (type == "4D_SU3_GAUGE_3x3") h->skip_third = 0;
  else r = MAG_ERR_IO;
  if (r || !aos) {
    std::fclose(f);
    return r;
  }
  const long nlink = 4 * volume(h->dims);
  const int per = h->skip_third ? 12 : 18;
  std::vector<double> data((size_t)nlink * per);
  const bool full = std::fread(data.data(), 8, data.size(), f) == data.size();
  std::fclose(f);
  if (!full) return MAG_ERR_IO;
  swap64((unsigned char*)data.data(), data.size());
  h->checksum_swapped = 0;
  if (checksum32(data.data(), data.size()) != h->checksum) {
    // 4D_SU3_GAUGE as the reference's export writes it (sum after the swap)
    if (!h->skip_third || checksum32_swapped(data.data(), data.size()) != h->checksum) return MAG_ERR_CHECKSUM;
    h->checksum_swapped = 1;
  }
  for (long l = 0; l < nlink; ++l) {
    double* d = aos + (size_t)l * 18;
    std::memcpy(d, &data[(size_t)l * per], per * sizeof(double));
    if (per == 12) std::memset(d + 12, 0, 6 * sizeof(double));
  }
  return MAG_OK;
}

int mag_export_config_nersc(mag_field U, const char* path, double beta, int sequence_number,
                            const char* ensemble_label, int skip_third) {
  if (!U || !path) return MAG_ERR_ARG;
  mag_nersc_header h{};
  int kind = 0, prec = 0;
  size_t bytes = 0;
  int r = mag_field_info(U, h.dims, &kind, &prec, &bytes);
  if (r) return r;
  if (kind != MAG_LINK18 && kind != MAG_LINK12) return fail(U->ctx, MAG_ERR_ARG, "export_config_nersc needs a link field");
  std::vector<double> aos((size_t)4 * volume(h.dims) * 18);
  if ((r = mag_field_download_aos(U, aos.data())) || (r = mag_avg_plaquette(U, &h.plaquette))) return r;
  double tr = 0.;  // site order, mu order (field-matrix.h:391-405)
  for (size_t l = 0; l < aos.size() / 18; ++l) tr += aos[l * 18 + 0] + aos[l * 18 + 8] + aos[l * 18 + 16];
  h.link_trace = tr / (12. * (double)volume(h.dims));
  h.skip_third = skip_third ? 1 : 0;
  h.beta = beta;
  h.sequence_number = sequence_number;
  std::snprintf(h.ensemble_label, sizeof h.ensemble_label, "%s", ensemble_label ? ensemble_label : "");
  r = mag_nersc_write(path, aos.data(), &h);
  if (r == MAG_ERR_IO) return fail(U->ctx, r, std::string("cannot write ") + path);
  return r;
}

int mag_import_config_nersc(mag_field U, const char* path, mag_nersc_header* hout) {
  if (!U || !path) return MAG_ERR_ARG;
  int dims[4], kind = 0, prec = 0;
  size_t bytes = 0;
  int r = mag_field_info(U, dims, &kind, &prec, &bytes);
  if (r) return r;
  if (kind != MAG_LINK18 && kind != MAG_LINK12) return fail(U->ctx, MAG_ERR_ARG, "import_config_nersc needs a link field");
  mag_nersc_header h{};
  if ((r = mag_nersc_read(path, nullptr, &h))) return fail(U->ctx, r, std::string("cannot read header of ") + path);
  for (int i = 0; i < 4; ++i)
    if (h.dims[i] != dims[i]) return fail(U->ctx, MAG_ERR_GEOMETRY, "lattice size differs from the file's");
  std::vector<double> aos((size_t)4 * volume(dims) * 18);
  if ((r = mag_nersc_read(path, aos.data(), &h)))
    return fail(U->ctx, r, r == MAG_ERR_CHECKSUM ? "checksum differs from the header's" : "cannot read links");
  if ((r = mag_field_upload_aos(U, aos.data()))) return r;
  if (h.skip_third && (r = mag_reunitarize(U, nullptr))) return r;
  double plaq = 0.;
  if ((r = mag_avg_plaquette(U, &plaq))) return r;
  if (hout) *hout = h;
  if (std::fabs(plaq - h.plaquette) > 1e-4)
    return fail(U->ctx, MAG_ERR_CHECKSUM, "plaquette differs from the header's by more than 1e-4");
  return MAG_OK;
}

}  // extern "C"
