// NERSC gauge configuration files (host side): the reader of
// ReadLatticeParallel / ReadLatticeSerial (ReadLatticePar.C:60-234) and the
// matching writer (WriteLatticePar.C, LatticeHeader::write latheader.C:
// 210-265), producing and taking the CANONICAL gauge field
// [t][z][y][x][mu][row][col][re,im] in FP64, ready for
// cps_hip_gauge_import_canonical.
//
// File format: an ASCII header of "KEY = value" lines from BEGIN_HEADER to
// END_HEADER, then per site (x fastest, then y, z, t) the 4 links mu = 0..3,
// each 3x3 (DATATYPE 4D_SU3_GAUGE_3x3) or its first two rows
// (4D_SU3_GAUGE; the third row is rebuilt as conj(row0 x row1),
// ReadLatticePar.C:178-193), complex numbers in the FLOATING_POINT format.
// Checks: CHECKSUM = the 32-bit sum of the data words read in the file's
// byte order (FPConv::checksum, fpconv.C:320-361); PLAQUETTE =
// sum ReTr U_munu / (18 V), compared with relative precision chkprec
// (ReadLatticePar.C:237-286, default 0.01); LINK_TRACE = sum ReTr U / (12 V)
// is reported but, as in the reference, not enforced.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cps_hip.h"

namespace {

typedef std::complex<double> zd;

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

bool host_big_endian() {
  const uint32_t one = 1;
  unsigned char c;
  std::memcpy(&c, &one, 1);
  return c == 0;
}

uint32_t bswap32(uint32_t v) {
  return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}

struct Header {
  std::map<std::string, std::string> kv;
  std::streamoff data_start = 0;
  bool get(const std::string& k, std::string* v) const {
    auto it = kv.find(k);
    if (it == kv.end()) return false;
    *v = it->second;
    return true;
  }
};

int read_header(std::ifstream& f, Header* h) {
  std::string line;
  bool begun = false;
  while (std::getline(f, line)) {
    const std::string t = trim(line);
    if (t == "BEGIN_HEADER") {
      begun = true;
      continue;
    }
    if (t == "END_HEADER") {
      if (!begun) return CPS_HIP_ERR_IO;
      h->data_start = f.tellg();
      return 0;
    }
    const size_t eq = line.find('=');
    if (eq != std::string::npos && eq > 0) h->kv[trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
  }
  return CPS_HIP_ERR_IO;
}

int parse_info(const Header& h, cps_hip_nersc_info* info) {
  std::memset(info, 0, sizeof(*info));
  std::string v;
  for (int d = 0; d < 4; d++) {
    if (!h.get("DIMENSION_" + std::to_string(d + 1), &v)) return CPS_HIP_ERR_IO;
    info->dims[d] = std::atoi(v.c_str());
    if (info->dims[d] <= 0) return CPS_HIP_ERR_IO;
  }
  if (!h.get("DATATYPE", &v)) return CPS_HIP_ERR_IO;
  if (v == "4D_SU3_GAUGE") info->recon_row_3 = 1;
  else if (v == "4D_SU3_GAUGE_3x3") info->recon_row_3 = 0;
  else return CPS_HIP_ERR_IO;
  if (!h.get("FLOATING_POINT", &v)) return CPS_HIP_ERR_IO;
  // FP_FORMAT_NAME of fpconv.C:12-22; "IEEE32"/"IEEE64" are big endian
  if (v == "IEEE32BIG" || v == "IEEE32") { info->bytes_per_real = 4; info->big_endian = 1; }
  else if (v == "IEEE32LITTLE") { info->bytes_per_real = 4; info->big_endian = 0; }
  else if (v == "IEEE64BIG" || v == "IEEE64") { info->bytes_per_real = 8; info->big_endian = 1; }
  else if (v == "IEEE64LITTLE") { info->bytes_per_real = 8; info->big_endian = 0; }
  else return CPS_HIP_ERR_IO;
  if (h.get("CHECKSUM", &v)) info->checksum_header = (unsigned int)std::strtoul(v.c_str(), nullptr, 16);
  if (h.get("PLAQUETTE", &v)) info->plaquette_header = std::atof(v.c_str());
  if (h.get("LINK_TRACE", &v)) info->link_trace_header = std::atof(v.c_str());
  return 0;
}

// U_mu(x) as a 3x3 complex matrix of the CANONICAL field
void load_link(const double* u, const int L[4], const int x[4], int mu, zd m[9]) {
  const int64_t site = x[0] + (int64_t)L[0] * (x[1] + (int64_t)L[1] * (x[2] + (int64_t)L[2] * x[3]));
  const double* p = u + 72 * site + 18 * mu;
  for (int k = 0; k < 9; k++) m[k] = zd(p[2 * k], p[2 * k + 1]);
}

void mul(const zd a[9], const zd b[9], zd c[9]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) c[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}

void mul_dag(const zd a[9], const zd b[9], zd c[9]) {   // a * b^dagger
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      c[3 * i + j] = a[3 * i] * std::conj(b[3 * j]) + a[3 * i + 1] * std::conj(b[3 * j + 1]) +
                     a[3 * i + 2] * std::conj(b[3 * j + 2]);
}

// sum over sites and mu < nu of ReTr U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+,
// / (18 V) (Lattice::SumReTrPlaq / 18 / VolSites, ReadLatticePar.C:242),
// and sum ReTr U / (12 V) (ReadLatticePar.C:260-269)
void plaq_linktrace(const double* u, const int L[4], double* plaq, double* ltrace) {
  const int64_t V = (int64_t)L[0] * L[1] * L[2] * L[3];
  double sp = 0, sl = 0;
  int x[4];
  for (x[3] = 0; x[3] < L[3]; x[3]++)
    for (x[2] = 0; x[2] < L[2]; x[2]++)
      for (x[1] = 0; x[1] < L[1]; x[1]++)
        for (x[0] = 0; x[0] < L[0]; x[0]++) {
          zd U[4][9];
          for (int mu = 0; mu < 4; mu++) {
            load_link(u, L, x, mu, U[mu]);
            sl += U[mu][0].real() + U[mu][4].real() + U[mu][8].real();
          }
          for (int mu = 0; mu < 4; mu++)
            for (int nu = mu + 1; nu < 4; nu++) {
              int xm[4] = {x[0], x[1], x[2], x[3]}, xn[4] = {x[0], x[1], x[2], x[3]};
              xm[mu] = (xm[mu] + 1) % L[mu];
              xn[nu] = (xn[nu] + 1) % L[nu];
              zd a[9], b[9], t1[9], t2[9], t3[9];
              load_link(u, L, xm, nu, a);
              load_link(u, L, xn, mu, b);
              mul(U[mu], a, t1);
              mul_dag(t1, b, t2);
              mul_dag(t2, U[nu], t3);
              sp += t3[0].real() + t3[4].real() + t3[8].real();
            }
        }
  *plaq = sp / (18.0 * V);
  *ltrace = sl / (12.0 * V);
}

// the 32-bit words of n file bytes as the file's byte order reads them
uint32_t checksum_words(const unsigned char* b, size_t n, bool file_big) {
  uint32_t s = 0;
  const bool swap = file_big != host_big_endian();
  for (size_t k = 0; k + 4 <= n; k += 4) {
    uint32_t w;
    std::memcpy(&w, b + k, 4);
    s += swap ? bswap32(w) : w;
  }
  return s;
}

double read_real(const unsigned char* b, int bytes, bool file_big) {
  unsigned char t[8];
  for (int k = 0; k < bytes; k++) t[k] = b[(file_big != host_big_endian()) ? bytes - 1 - k : k];
  if (bytes == 4) {
    float f;
    std::memcpy(&f, t, 4);
    return f;
  }
  double d;
  std::memcpy(&d, t, 8);
  return d;
}

void write_real(unsigned char* b, double v, int bytes, bool file_big) {
  unsigned char t[8];
  if (bytes == 4) {
    const float f = (float)v;
    std::memcpy(t, &f, 4);
  } else {
    std::memcpy(t, &v, 8);
  }
  for (int k = 0; k < bytes; k++) b[(file_big != host_big_endian()) ? bytes - 1 - k : k] = t[k];
}

}  // namespace

extern "C" {

int cps_hip_nersc_read_header(const char* path, cps_hip_nersc_info* info) {
  if (!path || !info) return CPS_HIP_ERR_ARG;
  std::ifstream f(path, std::ios::binary);
  if (!f.good()) return CPS_HIP_ERR_IO;
  Header h;
  int rc = read_header(f, &h);
  if (rc) return rc;
  return parse_info(h, info);
}

int cps_hip_nersc_read(const char* path, double* u_canon, cps_hip_nersc_info* info, double chkprec) {
  if (!path || !u_canon || !info) return CPS_HIP_ERR_ARG;
  std::ifstream f(path, std::ios::binary);
  if (!f.good()) return CPS_HIP_ERR_IO;
  Header h;
  int rc = read_header(f, &h);
  if (rc) return rc;
  if ((rc = parse_info(h, info))) return rc;
  const int* L = info->dims;
  const int64_t V = (int64_t)L[0] * L[1] * L[2] * L[3];
  const int per_link = info->recon_row_3 ? 12 : 18;
  const int bpr = info->bytes_per_real;
  const size_t site_bytes = (size_t)4 * per_link * bpr;
  std::vector<unsigned char> buf(site_bytes);
  uint32_t csum = 0;
  f.seekg(h.data_start);
  for (int64_t s = 0; s < V; s++) {
    f.read((char*)buf.data(), (std::streamsize)site_bytes);
    if (!f.good()) return CPS_HIP_ERR_IO;
    csum += checksum_words(buf.data(), site_bytes, info->big_endian != 0);
    for (int mu = 0; mu < 4; mu++) {
      double* m = u_canon + 72 * s + 18 * mu;
      for (int k = 0; k < per_link; k++)
        m[k] = read_real(buf.data() + (size_t)(mu * per_link + k) * bpr, bpr, info->big_endian != 0);
      if (info->recon_row_3) {
        // third row = conj(row0 x row1)
        m[12] = m[2] * m[10] - m[3] * m[11] - m[4] * m[8] + m[5] * m[9];
        m[13] = -m[2] * m[11] - m[3] * m[10] + m[4] * m[9] + m[5] * m[8];
        m[14] = -m[0] * m[10] + m[1] * m[11] + m[4] * m[6] - m[5] * m[7];
        m[15] = m[0] * m[11] + m[1] * m[10] - m[4] * m[7] - m[5] * m[6];
        m[16] = m[0] * m[8] - m[1] * m[9] - m[2] * m[6] + m[3] * m[7];
        m[17] = -m[0] * m[9] - m[1] * m[8] + m[2] * m[7] + m[3] * m[6];
      }
    }
  }
  info->checksum_data = csum;
  plaq_linktrace(u_canon, L, &info->plaquette_data, &info->link_trace_data);
  if (csum != info->checksum_header) return CPS_HIP_ERR_CHECKSUM;
  if (std::fabs((info->plaquette_data - info->plaquette_header) / info->plaquette_data) > chkprec)
    return CPS_HIP_ERR_PLAQUETTE;
  return 0;
}

int cps_hip_nersc_write(const char* path, const int dims[4], const double* u_canon, int recon_row_3,
                        int bytes_per_real, const char* ensemble_id, int sequence_number) {
  if (!path || !dims || !u_canon || (bytes_per_real != 4 && bytes_per_real != 8)) return CPS_HIP_ERR_ARG;
  for (int d = 0; d < 4; d++)
    if (dims[d] <= 0) return CPS_HIP_ERR_ARG;
  const int64_t V = (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
  const int per_link = recon_row_3 ? 12 : 18;
  const size_t site_bytes = (size_t)4 * per_link * bytes_per_real;
  std::vector<unsigned char> data((size_t)V * site_bytes);
  for (int64_t s = 0; s < V; s++)
    for (int mu = 0; mu < 4; mu++)
      for (int k = 0; k < per_link; k++)
        write_real(data.data() + (size_t)s * site_bytes + (size_t)(mu * per_link + k) * bytes_per_real,
                   u_canon[72 * s + 18 * mu + k], bytes_per_real, true);
  const uint32_t csum = checksum_words(data.data(), data.size(), true);
  double plaq, ltr;
  plaq_linktrace(u_canon, dims, &plaq, &ltr);
  std::ofstream f(path, std::ios::binary);
  if (!f.good()) return CPS_HIP_ERR_IO;
  char cs[16];
  std::snprintf(cs, sizeof(cs), "%08x", csum);
  std::ostringstream hd;
  hd.precision(10);
  hd << "BEGIN_HEADER\n"
     << "HDR_VERSION = 1.0\n"
     << "DATATYPE = " << (recon_row_3 ? "4D_SU3_GAUGE" : "4D_SU3_GAUGE_3x3") << "\n"
     << "STORAGE_FORMAT = 1.0\n";
  for (int d = 0; d < 4; d++) hd << "DIMENSION_" << d + 1 << " = " << dims[d] << "\n";
  hd << "LINK_TRACE = " << ltr << "\n"
     << "PLAQUETTE = " << plaq << "\n";
  for (int d = 0; d < 4; d++) hd << "BOUNDARY_" << d + 1 << " = PERIODIC\n";
  hd << "CHECKSUM = " << cs << "\n"
     << "ENSEMBLE_ID = " << (ensemble_id ? ensemble_id : "") << "\n"
     << "ENSEMBLE_LABEL = \n"
     << "SEQUENCE_NUMBER = " << sequence_number << "\n"
     << "CREATOR = cps_hip\n"
     << "CREATOR_HARDWARE = \n"
     << "CREATION_DATE = \n"
     << "ARCHIVE_DATE = \n"
     << "FLOATING_POINT = " << (bytes_per_real == 4 ? "IEEE32BIG" : "IEEE64BIG") << "\n"
     << "END_HEADER\n";
  const std::string hs = hd.str();
  f.write(hs.data(), (std::streamsize)hs.size());
  f.write((const char*)data.data(), (std::streamsize)data.size());
  return f.good() ? 0 : CPS_HIP_ERR_IO;
}

}  // extern "C"
