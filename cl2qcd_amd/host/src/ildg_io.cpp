// ILDG/LIME reading and writing of gauge configurations, host side.
//
// A LIME record is a 144-byte header -- magic 0x456789ab, version 1, flags (message begin 0x8000, message
// end 0x4000), 64-bit payload length, 128-byte type string, all big-endian -- followed by the payload padded
// to 8 bytes.  The ildg-binary-data payload holds [t][z][y][x][mu_ildg][3][3] big-endian complex numbers
// with mu_ildg = x, y, z, t; internally mu = T, X, Y, Z, so internal mu = (mu_ildg + 1) % 4.
#include "../include/ildg_io/ildgIo.hpp"

#include <cerrno>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iterator>
#include <sstream>
#include <stdexcept>
#include <sys/stat.h>

namespace ildgIo {

namespace {

constexpr uint32_t LIME_MAGIC = 0x456789abu;
constexpr uint16_t LIME_MB = 0x8000, LIME_ME = 0x4000;

uint32_t crc_table_entry(uint32_t c)
{
    for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xedb88320u ^ (c >> 1) : c >> 1;
    return c;
}

struct CrcTable {
    uint32_t t[256];
    CrcTable()
    {
        for (uint32_t n = 0; n < 256; ++n) t[n] = crc_table_entry(n);
    }
};

uint32_t rotl(uint32_t v, uint32_t r)
{
    return r == 0 ? v : (v << r) | (v >> (32 - r));
}

uint64_t get_be(const unsigned char* p, int n)
{
    uint64_t v = 0;
    for (int i = 0; i < n; ++i) v = (v << 8) | p[i];
    return v;
}

void put_be(std::string& out, uint64_t v, int n)
{
    for (int i = n - 1; i >= 0; --i) out.push_back((char)((v >> (8 * i)) & 0xff));
}

void lime_record(std::string& out, const std::string& type, const char* data, size_t len, uint16_t flags)
{
    put_be(out, LIME_MAGIC, 4);
    put_be(out, 1, 2);
    put_be(out, flags, 2);
    put_be(out, len, 8);
    std::string t(128, '\0');
    std::memcpy(&t[0], type.data(), std::min<size_t>(type.size(), 127));
    out += t;
    out.append(data, len);
    out.append((8 - len % 8) % 8, '\0');
}

int xml_int(const std::string& xml, const std::string& tag)
{
    const size_t q = xml.find("<" + tag + ">");
    return q == std::string::npos ? -1 : std::atoi(xml.c_str() + q + tag.size() + 2);
}

uint32_t xml_hex(const std::string& xml, const std::string& tag, bool* found)
{
    const size_t q = xml.find("<" + tag + ">");
    *found         = q != std::string::npos;
    return *found ? (uint32_t)std::strtoul(xml.c_str() + q + tag.size() + 2, nullptr, 16) : 0u;
}

/* shortest decimal form that reads back to the same double */
std::string num(double v)
{
    char b[32];
    const auto r = std::to_chars(b, b + sizeof b, v);
    return std::string(b, r.ptr);
}

bool file_exists(const std::string& f)
{
    struct stat st;
    return ::stat(f.c_str(), &st) == 0;
}

}  // namespace

uint32_t crc32(const unsigned char* p, size_t n)
{
    static const CrcTable table;
    uint32_t c = 0xffffffffu;
    for (size_t i = 0; i < n; ++i) c = table.t[(c ^ p[i]) & 0xffu] ^ (c >> 8);
    return c ^ 0xffffffffu;
}

Checksum calculate_ildg_checksum(const char* buf, size_t nbytes, size_t NT, size_t NS)
{
    const size_t vol = NT * NS * NS * NS;
    if (vol == 0 || nbytes % vol != 0 || (nbytes / vol != 4 * 18 * 8 && nbytes / vol != 4 * 18 * 4))
        throw std::invalid_argument("calculate_ildg_checksum: buffer does not hold a gauge field of this size");
    const size_t per_site = nbytes / vol;
    Checksum s;
    for (size_t rank = 0; rank < vol; ++rank) {  // file order t, z, y, x is the rank order
        const uint32_t w = crc32(reinterpret_cast<const unsigned char*>(buf) + rank * per_site, per_site);
        s.suma ^= rotl(w, (uint32_t)(rank % 29));
        s.sumb ^= rotl(w, (uint32_t)(rank % 31));
    }
    return s;
}

std::vector<Matrixsu3> readGaugefieldFromSourcefile(const std::string& file, int ns, int nt,
                                                    Sourcefileparameters* info, bool ignore_checksum)
{
    std::ifstream in(file, std::ios::binary);
    if (!in) throw std::invalid_argument("cannot open gauge configuration " + file);
    const std::vector<unsigned char> d((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    Sourcefileparameters p;
    size_t off = 0, data = 0, len_data = 0;
    bool have_format = false, have_data = false;
    while (off + 144 <= d.size()) {
        if (get_be(&d[off], 4) != LIME_MAGIC) throw std::runtime_error(file + ": not a LIME file");
        const uint64_t len = get_be(&d[off + 8], 8);
        const std::string type(reinterpret_cast<const char*>(&d[off + 16]),
                               strnlen(reinterpret_cast<const char*>(&d[off + 16]), 128));
        const size_t payload = off + 144;
        if (len > d.size() || payload + len > d.size()) throw std::runtime_error(file + ": truncated LIME record");
        const std::string text(reinterpret_cast<const char*>(&d[payload]), type == "ildg-binary-data" ? 0 : len);
        if (type == "ildg-format") {
            p.precision = xml_int(text, "precision");
            p.lx        = xml_int(text, "lx");
            p.ly        = xml_int(text, "ly");
            p.lz        = xml_int(text, "lz");
            p.lt        = xml_int(text, "lt");
            have_format = true;
        } else if (type == "ildg-binary-data") {
            data      = payload;
            len_data  = len;
            have_data = true;
        } else if (type == "scidac-checksum") {
            bool fa = false, fb = false;
            p.checksum.suma = xml_hex(text, "suma", &fa);
            p.checksum.sumb = xml_hex(text, "sumb", &fb);
            p.has_checksum  = fa && fb;
        } else if (type == "xlf-info") {
            p.xlf_info = text;
        }
        off = payload + ((len + 7) / 8) * 8;
    }
    if (!have_format || !have_data) throw std::runtime_error(file + ": no ildg-format or ildg-binary-data record");
    if (p.lx != ns || p.ly != ns || p.lz != ns || p.lt != nt)
        throw std::invalid_argument(file + ": lattice extents differ from the run parameters");
    if (p.precision != 64 && p.precision != 32) throw std::runtime_error(file + ": precision must be 32 or 64");
    const size_t vol = (size_t)ns * ns * ns * nt, esz = p.precision == 32 ? 4 : 8;
    if (len_data != vol * 4 * 18 * esz) throw std::runtime_error(file + ": ildg-binary-data has the wrong size");
    if (p.has_checksum && !ignore_checksum) {
        const Checksum c = calculate_ildg_checksum(reinterpret_cast<const char*>(&d[data]), len_data, nt, ns);
        if (c != p.checksum) throw std::runtime_error(file + ": checksum of the links differs from the file's");
    }
    std::vector<Matrixsu3> links(vol * 4);
    double* host = reinterpret_cast<double*>(links.data());
    for (size_t s = 0; s < vol; ++s)
        for (int l = 0; l < 4; ++l)
            for (int k = 0; k < 18; ++k) {
                const unsigned char* src = &d[data + ((s * 4 + l) * 18 + k) * esz];
                double v;
                if (esz == 8) {
                    const uint64_t b = get_be(src, 8);
                    std::memcpy(&v, &b, 8);
                } else {
                    const uint32_t b = (uint32_t)get_be(src, 4);
                    float f;
                    std::memcpy(&f, &b, 4);
                    v = f;
                }
                host[((l + 1) % 4 + 4 * s) * 18 + k] = v;
            }
    if (info) *info = p;
    return links;
}

void writeGaugefieldToFile(const std::string& file, const std::vector<Matrixsu3>& links, int ns, int nt,
                           const XlfInfo& info, int precision)
{
    if (precision != 64 && precision != 32) throw std::invalid_argument("writeGaugefieldToFile: precision 32 or 64");
    const size_t vol = (size_t)ns * ns * ns * nt, esz = precision == 32 ? 4 : 8;
    if (links.size() != 4 * vol) throw std::invalid_argument("writeGaugefieldToFile: links of the wrong size");
    const double* host = reinterpret_cast<const double*>(links.data());
    std::string bin;
    bin.reserve(vol * 4 * 18 * esz);
    for (size_t s = 0; s < vol; ++s)
        for (int l = 0; l < 4; ++l)
            for (int k = 0; k < 18; ++k) {
                const double v = host[((l + 1) % 4 + 4 * s) * 18 + k];
                if (esz == 8) {
                    uint64_t b;
                    std::memcpy(&b, &v, 8);
                    put_be(bin, b, 8);
                } else {
                    const float f = (float)v;
                    uint32_t b;
                    std::memcpy(&b, &f, 4);
                    put_be(bin, b, 4);
                }
            }
    const Checksum sum = calculate_ildg_checksum(bin.data(), bin.size(), nt, ns);
    const time_t now   = std::time(nullptr);
    char date[64]      = {0};
    std::strftime(date, sizeof date, "%a %b %d %H:%M:%S %Y", std::localtime(&now));
    std::ostringstream xlf;
    xlf << "plaquette = " << num(info.plaquette) << "\n"
        << "trajectory nr = " << info.trajectory << "\n"
        << "beta = " << num(info.beta) << "\n"
        << "kappa = " << num(info.kappa) << "\n"
        << "mu = " << num(info.mu) << "\n"
        << "c2_rec = 0\n"
        << "time = " << (long long)now << "\n"
        << "hmcversion = cl2qcd_amd\n"
        << "mubar = " << num(info.mubar) << "\n"
        << "epsilonbar = 0\n"
        << "date = " << date << "\n";
    std::ostringstream fmt;
    fmt << "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n<ildgFormat xmlns=\"http://www.lqcd.org/ildg\"\n"
           "            xmlns:xsi=\"http://www.w3.org/2001/XMLSchema-instance\"\n"
           "            xsi:schemaLocation=\"http://www.lqcd.org/ildg filefmt.xsd\">\n  <version>1.0</version>\n"
        << "  <field>su3gauge</field>\n  <precision>" << precision << "</precision>\n  <lx>" << ns << "</lx>\n  <ly>"
        << ns << "</ly>\n  <lz>" << ns << "</lz>\n  <lt>" << nt << "</lt>\n</ildgFormat>";
    std::ostringstream chk;
    chk << "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n<scidacChecksum>\n<version>1.0</version>\n<suma>" << std::hex
        << sum.suma << "</suma>\n<sumb>" << sum.sumb << "</sumb>\n</scidacChecksum>";
    // every record is its own LIME message; the first also carries the message-end flag, as the reference's
    // writer leaves it (its end flag counts up from 1 and only 1 sets the bit)
    std::string out;
    const std::string x = xlf.str(), f = fmt.str(), c = chk.str();
    lime_record(out, "xlf-info", x.data(), x.size(), LIME_MB | LIME_ME);
    lime_record(out, "ildg-format", f.data(), f.size(), LIME_MB);
    lime_record(out, "ildg-binary-data", bin.data(), bin.size(), LIME_MB);
    lime_record(out, "scidac-checksum", c.data(), c.size(), LIME_MB);
    if (file_exists(file)) {
        const std::string backup = file + "_backup";
        if (std::rename(file.c_str(), backup.c_str()) != 0)
            throw std::runtime_error("cannot move the existing " + file + " to " + backup + ": " + std::strerror(errno));
    }
    {
        std::ofstream o(file, std::ios::binary | std::ios::trunc);
        if (!o) throw std::runtime_error("cannot write " + file);
        o.write(out.data(), (std::streamsize)out.size());
        if (!o) throw std::runtime_error("writing " + file + " failed");
    }
    // read back, as the reference does after writing
    Sourcefileparameters back;
    (void)readGaugefieldFromSourcefile(file, ns, nt, &back);
    if (!back.has_checksum || back.checksum != sum) throw std::runtime_error(file + ": written checksum does not read back");
}

}  // namespace ildgIo
