// NERSC gauge-configuration reader: the host side of
// ReadLatticeParallel::read (src/util/qcdio/comsrc/ReadLatticePar.C:26-205)
// for one process.  Header as LatticeHeader::read (latheader.C:267-312),
// floating-point formats and the 32-bit checksum as FPConv
// (fpconv.C:12-22, 320-360), data order and row-3 reconstruction as
// SerialIO::load (iostyle.C:722-840) and ReadLatticePar.C:159-175.
// The plaquette / link-trace check (ReadLatticePar.C:396-450) runs on the
// device through dwfgpu_plaquette once the operator holds the links.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "dwf_gpu.h"

namespace {

// leading and trailing blanks removed (latheader.C:30-36 trims ' ' only)
std::string trim_blanks(const std::string& s) {
  const size_t a = s.find_first_not_of(' ');
  if (a == std::string::npos) return s;
  const size_t b = s.find_last_not_of(' ');
  return s.substr(a, b - a + 1);
}

struct Format {
  int bytes;  // 4 or 8, 0 unknown
  bool big;
};

// FP_FORMAT_NAME (fpconv.C:12-22); "AUTOMATIC" is the host format, the
// unsuffixed names are big-endian (fpconv.C:379-392)
Format parse_format(const std::string& name) {
  if (name == "IEEE32" || name == "IEEE32BIG") return {4, true};
  if (name == "IEEE32LITTLE") return {4, false};
  if (name == "IEEE64" || name == "IEEE64BIG") return {8, true};
  if (name == "IEEE64LITTLE" || name == "AUTOMATIC") return {8, false};
  return {0, false};  // TIDSP32 and anything else: not readable here
}

uint32_t load_u32(const unsigned char* p, bool big) {
  return big ? (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]
             : (uint32_t)p[3] << 24 | (uint32_t)p[2] << 16 | (uint32_t)p[1] << 8 | p[0];
}

double load_real(const unsigned char* p, const Format& f) {
  if (f.bytes == 4) {
    const uint32_t w = load_u32(p, f.big);
    float v;
    std::memcpy(&v, &w, 4);
    return v;
  }
  const uint64_t hi = load_u32(p + (f.big ? 0 : 4), f.big);
  const uint64_t lo = load_u32(p + (f.big ? 4 : 0), f.big);
  const uint64_t w = hi << 32 | lo;
  double v;
  std::memcpy(&v, &w, 8);
  return v;
}

}  // namespace

extern "C" int dwfgpu_nersc_read(const char* path, double* gauge, size_t n_reals,
                                 dwfgpu_nersc_header* hdr) {
  if (!path || !hdr) return DWFGPU_BAD_ARG;
  std::memset(hdr, 0, sizeof(*hdr));
  std::ifstream in(path, std::ios::binary);
  if (!in.good()) return DWFGPU_IO_ERROR;
  // key = value lines up to END_HEADER; the first occurrence of a key wins
  std::map<std::string, std::string> kv;
  std::string line;
  bool end = false;
  while (std::getline(in, line)) {
    if (line.find("END_HEADER") != std::string::npos) {
      end = true;
      break;
    }
    const size_t eq = line.find('=');
    if (eq != std::string::npos && eq > 0)
      kv.insert({trim_blanks(line.substr(0, eq)), trim_blanks(line.substr(eq + 1))});
  }
  if (!end) return DWFGPU_IO_ERROR;
  hdr->data_start = (long long)in.tellg();
  auto get = [&](const char* k) -> const std::string* {
    auto it = kv.find(k);
    return it == kv.end() ? nullptr : &it->second;
  };
  const std::string* dt = get("DATATYPE");
  hdr->recon_row_3 = dt && *dt == "4D_SU3_GAUGE" ? 1 : 0;
  const char* dk[4] = {"DIMENSION_1", "DIMENSION_2", "DIMENSION_3", "DIMENSION_4"};
  for (int d = 0; d < 4; ++d) {
    const std::string* v = get(dk[d]);
    if (!v || std::sscanf(v->c_str(), "%d", &hdr->dims[d]) != 1 || hdr->dims[d] <= 0)
      return DWFGPU_IO_ERROR;
  }
  // LINK_TRACE / PLAQUETTE are read as single precision (latheader.C:121-136)
  float f = 0.0f;
  if (const std::string* v = get("LINK_TRACE"))
    if (std::sscanf(v->c_str(), "%f", &f) == 1) hdr->link_trace = f;
  f = 0.0f;
  if (const std::string* v = get("PLAQUETTE"))
    if (std::sscanf(v->c_str(), "%f", &f) == 1) hdr->plaquette = f;
  unsigned int cs = 0;
  if (const std::string* v = get("CHECKSUM")) std::sscanf(v->c_str(), "%x", &cs);
  hdr->checksum = cs;
  const std::string* fp = get("FLOATING_POINT");
  const Format fmt = parse_format(fp ? *fp : std::string());
  if (!fmt.bytes) return DWFGPU_IO_ERROR;
  hdr->fp_bytes = fmt.bytes;
  hdr->big_endian = fmt.big ? 1 : 0;
  if (!gauge) return DWFGPU_OK;

  const size_t v4 = (size_t)hdr->dims[0] * hdr->dims[1] * hdr->dims[2] * hdr->dims[3];
  if (n_reals != v4 * 4 * 18) return DWFGPU_BAD_ARG;
  const int per_mat = hdr->recon_row_3 ? 12 : 18;
  const size_t site_bytes = (size_t)4 * per_mat * fmt.bytes;
  std::vector<unsigned char> buf(site_bytes);
  uint32_t sum = 0;
  // file order: t, z, y, x (x fastest), then mu, then the matrix row-major
  // (iostyle.C:786-812) -- the same as CANONICAL order
  for (size_t site = 0; site < v4; ++site) {
    in.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)site_bytes);
    if ((size_t)in.gcount() != site_bytes) return DWFGPU_IO_ERROR;
    // checksum: the data as 32-bit words in the file's byte order
    for (size_t w = 0; w < site_bytes; w += 4) sum += load_u32(buf.data() + w, fmt.big);
    for (int mu = 0; mu < 4; ++mu) {
      double* m = gauge + (site * 4 + mu) * 18;
      const unsigned char* src = buf.data() + (size_t)mu * per_mat * fmt.bytes;
      for (int k = 0; k < per_mat; ++k) m[k] = load_real(src + (size_t)k * fmt.bytes, fmt);
      if (hdr->recon_row_3) {
        // row 2 = conj(row 0 x row 1)  (ReadLatticePar.C:166-172)
        const double* a = m;      // row 0: a_j = (m[2j], m[2j+1])
        const double* b = m + 6;  // row 1
        for (int j = 0; j < 3; ++j) {
          const int p = (j + 1) % 3, q = (j + 2) % 3;
          const double re = a[2 * p] * b[2 * q] - a[2 * p + 1] * b[2 * q + 1] -
                            (a[2 * q] * b[2 * p] - a[2 * q + 1] * b[2 * p + 1]);
          const double im = a[2 * p] * b[2 * q + 1] + a[2 * p + 1] * b[2 * q] -
                            (a[2 * q] * b[2 * p + 1] + a[2 * q + 1] * b[2 * p]);
          m[12 + 2 * j] = re;
          m[12 + 2 * j + 1] = -im;
        }
      }
    }
  }
  hdr->checksum_data = sum;
  return sum == hdr->checksum ? DWFGPU_OK : DWFGPU_CHECKSUM_ERROR;
}
