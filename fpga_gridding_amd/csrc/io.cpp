// Plain host I/O around the gridder: OSKAR binary visibility files, FITS
// kernel cubes in, FITS grid cubes out.  Written from the formats, not from
// the reference's libraries (it uses its own oskar_binary sources and
// cfitsio):
//
// * OSKAR binary: a 64-byte header ("OSKARBIN\0", format version at byte 9)
//   followed by chunks of a 20-byte tag, optional extended names, the
//   payload and, from format version 2, an optional little-endian CRC-32C of
//   tag + names + payload (utils/binary/private_binary.h:40-160).  Payloads
//   are returned as stored, as the reference's reader does
//   (oskar_binary_read.c:39-106).
// * FITS (NASA FITS standard 4.0): 2880-byte blocks of 80-character header
//   cards, big-endian data.  The kernel reader takes what read_kernel reads
//   (read_kernel.c:35-118): a 3-D primary image (NAXIS1 = conv_size_half,
//   NAXIS3 = planes), the OVERSAMP, GRIDSIZE, CELLSIZE and W_SCALE keywords
//   and column 1 of the W_KERNELS binary table.  The writer writes what
//   write_fits_cube writes (write_fits_cube.c:37-168).
#include "fpga_gridding_amd_io.h"

#include <algorithm>
#include <cctype>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

namespace {

// OSKAR status codes used here (utils/oskar_global.h:43-64 and
// utils/binary/oskar_binary.h:96-117).
enum
{
    kErrFileIO = -2,
    kErrInvalidArgument = -3,
    kErrMemAlloc = -6,
    kBinOpenFail = -100,
    kBinSeekFail = -101,
    kBinReadFail = -102,
    kBinFileInvalid = -106,
    kBinFormatBad = -107,
    kBinVersionUnknown = -109,
    kBinMemoryNotAllocated = -114,
    kBinTagNotFound = -115,
    kBinTagTooLong = -116,
    kBinTagOutOfRange = -117,
    kBinCrcFail = -118
};

constexpr int kBinFormatVersion = 2;

// CRC-32C (Castagnoli, reflected polynomial 0x82F63B78), bytewise table.
struct Crc32c
{
    uint32_t t[256];
    Crc32c()
    {
        for (uint32_t i = 0; i < 256; ++i)
        {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c >> 1) ^ ((c & 1u) * 0x82F63B78u);
            t[i] = c;
        }
    }
    // Running register (start 0xFFFFFFFF; the code is the register ^ ~0).
    uint32_t update(uint32_t reg, const void* data, size_t n) const
    {
        const unsigned char* p = static_cast<const unsigned char*>(data);
        for (size_t i = 0; i < n; ++i) reg = (reg >> 8) ^ t[(reg ^ p[i]) & 0xFFu];
        return reg;
    }
};

const Crc32c& crc_table()
{
    static const Crc32c c;
    return c;
}

uint32_t le32(const unsigned char* p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
           ((uint32_t)p[3] << 24);
}

uint64_t le64(const unsigned char* p)
{
    return (uint64_t)le32(p) | ((uint64_t)le32(p + 4) << 32);
}

}  // namespace

struct gw_binary
{
    FILE* f = nullptr;
    int version = 0;
    int search_start = 0;     // first chunk a query looks at
    struct Chunk
    {
        bool extended;
        int data_type, id_group, id_tag, user_index;
        std::string name_group, name_tag;
        long long payload_offset;
        size_t payload_size;
        bool has_crc;
        uint32_t crc_reg;     // CRC register after the tag and names
        uint32_t crc_file;    // the stored code
    };
    std::vector<Chunk> chunks;
};

namespace {

// Index every chunk of an open file; returns 0 or an OSKAR status.
int index_binary(gw_binary& h)
{
    unsigned char hdr[64];
    if (std::fread(hdr, 1, 64, h.f) != 64) return kBinFileInvalid;
    if (std::memcmp(hdr, "OSKARBIN", 9) != 0) return kBinFileInvalid;
    h.version = hdr[9];
    if (h.version < 1 || h.version > kBinFormatVersion)
        return kBinVersionUnknown;
    if (h.version == 1 && hdr[10] != 0) return kBinFormatBad;  // big-endian
    const Crc32c& crc = crc_table();
    for (;;)
    {
        unsigned char tag[20];
        const size_t got = std::fread(tag, 1, 20, h.f);
        if (got == 0 && std::feof(h.f)) break;
        if (got != 20) return kBinFileInvalid;
        if (tag[0] != 'T' || tag[2] != 'G') return kBinFileInvalid;
        const int fv = tag[1] - 0x40;
        if (fv < 1 || fv > kBinFormatVersion) return kBinVersionUnknown;
        const unsigned char flags = tag[4];
        if (fv > 1 && (flags & 0x20)) return kBinFormatBad;  // big-endian
        gw_binary::Chunk c;
        c.extended = (flags & 0x80) != 0;
        c.data_type = tag[5];
        c.id_group = tag[6];
        c.id_tag = tag[7];
        c.user_index = (int)le32(tag + 8);
        const uint64_t block = le64(tag + 12);
        c.has_crc = fv > 1 && (flags & 0x40);
        uint64_t payload = block;
        uint32_t reg = crc.update(0xFFFFFFFFu, tag, 20);
        if (c.has_crc)
        {
            if (payload < 4) return kBinFileInvalid;
            payload -= 4;
        }
        if (c.extended)
        {
            const size_t lg = tag[6], lt = tag[7];
            if (payload < lg + lt) return kBinFileInvalid;
            payload -= lg + lt;
            std::vector<char> g(lg + 1, 0), t(lt + 1, 0);
            if ((lg && std::fread(g.data(), 1, lg, h.f) != lg) ||
                    (lt && std::fread(t.data(), 1, lt, h.f) != lt))
                return kBinFileInvalid;
            reg = crc.update(reg, g.data(), lg);
            reg = crc.update(reg, t.data(), lt);
            c.name_group = g.data();
            c.name_tag = t.data();
        }
        c.payload_offset = (long long)std::ftell(h.f);
        c.payload_size = (size_t)payload;
        c.crc_reg = reg;
        c.crc_file = 0;
        if (std::fseek(h.f, (long)payload, SEEK_CUR) != 0)
            return kBinFileInvalid;
        if (c.has_crc)
        {
            unsigned char code[4];
            if (std::fread(code, 1, 4, h.f) != 4) return kBinFileInvalid;
            c.crc_file = le32(code);
        }
        h.chunks.push_back(std::move(c));
    }
    return 0;
}

void read_chunk(gw_binary* h, int i, size_t data_size, void* data,
                int* status)
{
    if (*status) return;
    if (i < 0 || i >= (int)h->chunks.size())
    {
        *status = kBinTagOutOfRange;
        return;
    }
    const gw_binary::Chunk& c = h->chunks[i];
    if (c.payload_size == 0) return;
    if (!data || data_size < c.payload_size)
    {
        *status = kBinMemoryNotAllocated;
        return;
    }
    if (std::fseek(h->f, (long)c.payload_offset, SEEK_SET) != 0)
    {
        *status = kBinSeekFail;
        return;
    }
    if (std::fread(data, 1, c.payload_size, h->f) != c.payload_size)
    {
        *status = kBinReadFail;
        return;
    }
    if (c.has_crc)
    {
        const uint32_t code = crc_table().update(c.crc_reg, data,
                c.payload_size) ^ 0xFFFFFFFFu;
        if (code != c.crc_file) *status = kBinCrcFail;
    }
}

// ------------------------------------------------------------------------
// FITS
// ------------------------------------------------------------------------

constexpr size_t kBlock = 2880;

struct FitsCard
{
    std::string key, value;
};

// Header cards of the HDU at `pos` up to END; returns the offset of its
// data (block aligned) or 0 on a malformed header.
size_t read_header(const std::vector<unsigned char>& f, size_t pos,
                   std::vector<FitsCard>& cards)
{
    cards.clear();
    for (size_t at = pos; at + 80 <= f.size(); at += 80)
    {
        std::string card(reinterpret_cast<const char*>(&f[at]), 80);
        std::string key = card.substr(0, 8);
        key.erase(key.find_last_not_of(' ') + 1);
        if (key == "END")
            return ((at + 80 - pos + kBlock - 1) / kBlock) * kBlock + pos;
        if (card.size() >= 10 && card[8] == '=' && card[9] == ' ')
        {
            std::string v = card.substr(10);
            // String values are quoted; others end at a comment slash.
            size_t b = v.find_first_not_of(' ');
            if (b != std::string::npos && v[b] == '\'')
            {
                std::string s;
                size_t i = b + 1;
                for (; i < v.size(); ++i)
                {
                    if (v[i] == '\'')
                    {
                        if (i + 1 < v.size() && v[i + 1] == '\'')
                        {
                            s += '\'';
                            ++i;
                            continue;
                        }
                        break;
                    }
                    s += v[i];
                }
                s.erase(s.find_last_not_of(' ') + 1);
                v = s;
            }
            else
            {
                const size_t slash = v.find('/');
                if (slash != std::string::npos) v = v.substr(0, slash);
                v.erase(0, v.find_first_not_of(' '));
                v.erase(v.find_last_not_of(' ') + 1);
            }
            cards.push_back({key, v});
        }
    }
    return 0;
}

const std::string* find_key(const std::vector<FitsCard>& c,
                            const std::string& k)
{
    for (const FitsCard& x : c)
        if (x.key == k) return &x.value;
    return nullptr;
}

bool key_long(const std::vector<FitsCard>& c, const std::string& k,
              long long* out)
{
    const std::string* v = find_key(c, k);
    if (!v || v->empty()) return false;
    char* end = nullptr;
    const long long x = std::strtoll(v->c_str(), &end, 10);
    if (end == v->c_str()) return false;
    *out = x;
    return true;
}

bool key_double(const std::vector<FitsCard>& c, const std::string& k,
                double* out)
{
    const std::string* v = find_key(c, k);
    if (!v || v->empty()) return false;
    std::string s = *v;
    std::replace(s.begin(), s.end(), 'D', 'E');   // FITS exponent letter
    char* end = nullptr;
    const double x = std::strtod(s.c_str(), &end);
    if (end == s.c_str()) return false;
    *out = x;
    return true;
}

// Big-endian element i of a FITS array as double (BITPIX or TFORM type).
double be_value(const unsigned char* p, int bitpix, size_t i)
{
    switch (bitpix)
    {
    case 8:
        return (double)p[i];
    case 16:
    {
        const unsigned char* q = p + 2 * i;
        return (double)(int16_t)(((uint16_t)q[0] << 8) | q[1]);
    }
    case 32:
    {
        const unsigned char* q = p + 4 * i;
        return (double)(int32_t)(((uint32_t)q[0] << 24) |
                ((uint32_t)q[1] << 16) | ((uint32_t)q[2] << 8) | q[3]);
    }
    case 64:
    {
        const unsigned char* q = p + 8 * i;
        uint64_t u = 0;
        for (int k = 0; k < 8; ++k) u = (u << 8) | q[k];
        return (double)(int64_t)u;
    }
    case -32:
    {
        const unsigned char* q = p + 4 * i;
        const uint32_t u = ((uint32_t)q[0] << 24) | ((uint32_t)q[1] << 16) |
                ((uint32_t)q[2] << 8) | q[3];
        float f;
        std::memcpy(&f, &u, 4);
        return (double)f;
    }
    default:
    {
        const unsigned char* q = p + 8 * i;
        uint64_t u = 0;
        for (int k = 0; k < 8; ++k) u = (u << 8) | q[k];
        double d;
        std::memcpy(&d, &u, 8);
        return d;
    }
    }
}

bool read_file(const char* path, std::vector<unsigned char>& out)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    in.seekg(0, std::ios::end);
    const std::streamoff n = in.tellg();
    if (n < 0) return false;
    in.seekg(0);
    out.resize((size_t)n);
    return n == 0 || (bool)in.read(reinterpret_cast<char*>(out.data()), n);
}

void card(std::string& h, const std::string& key, const std::string& value,
          const std::string& comment = "")
{
    char buf[81];
    std::string v = value;
    if (v.size() < 20) v = std::string(20 - v.size(), ' ') + v;
    std::string line = key;
    line.resize(8, ' ');
    line += "= " + v;
    if (!comment.empty()) line += " / " + comment;
    line.resize(80, ' ');
    std::memcpy(buf, line.data(), 80);
    h.append(buf, 80);
}

// A new primary image of `bitpix` (-32 or -64) with dims, all zero.
bool create_fits(const std::string& path, int bitpix, long w, long hgt,
                 long planes)
{
    std::string h;
    card(h, "SIMPLE", "T", "file conforms to FITS standard");
    card(h, "BITPIX", std::to_string(bitpix), "number of bits per data pixel");
    card(h, "NAXIS", "3", "number of data axes");
    card(h, "NAXIS1", std::to_string(w), "length of data axis 1");
    card(h, "NAXIS2", std::to_string(hgt), "length of data axis 2");
    card(h, "NAXIS3", std::to_string(planes), "length of data axis 3");
    card(h, "EXTEND", "T", "FITS dataset may contain extensions");
    std::string end = "END";
    end.resize(80, ' ');
    h += end;
    h.resize(((h.size() + kBlock - 1) / kBlock) * kBlock, ' ');
    const size_t bytes = (size_t)w * hgt * planes * (bitpix == -64 ? 8 : 4);
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) return false;
    out.write(h.data(), (std::streamsize)h.size());
    std::vector<char> zero(kBlock, 0);
    size_t left = ((bytes + kBlock - 1) / kBlock) * kBlock;
    while (left)
    {
        const size_t n = std::min(left, kBlock);
        out.write(zero.data(), (std::streamsize)n);
        left -= n;
    }
    return (bool)out;
}

// Writes num values (float or double, `precision`) big-endian at element
// `first` of an existing floating-point primary image of `bitpix`,
// converting between float and double as needed.
bool write_values(const std::string& path, size_t data_off, int bitpix,
                  size_t first, size_t num, int precision, const void* data)
{
    std::fstream io(path, std::ios::binary | std::ios::in | std::ios::out);
    if (!io) return false;
    const size_t es = bitpix == -64 ? 8 : 4;
    std::vector<unsigned char> buf(num * es);
    for (size_t i = 0; i < num; ++i)
    {
        const double x = precision == GW_OSKAR_DOUBLE ?
                static_cast<const double*>(data)[i] :
                (double)static_cast<const float*>(data)[i];
        unsigned char le[8];
        if (es == 8)
            std::memcpy(le, &x, 8);
        else
        {
            const float f = (float)x;
            std::memcpy(le, &f, 4);
        }
        for (size_t k = 0; k < es; ++k)
            buf[i * es + k] = le[es - 1 - k];   // host is little-endian
    }
    io.seekp((std::streamoff)(data_off + first * es));
    io.write(reinterpret_cast<const char*>(buf.data()),
             (std::streamsize)buf.size());
    return (bool)io;
}

// The primary image geometry of an existing FITS file, if it is 3-D.
bool image_geometry(const std::string& path, int* bitpix, long long dims[3],
                    size_t* data_off)
{
    std::vector<unsigned char> f;
    if (!read_file(path.c_str(), f)) return false;
    std::vector<FitsCard> c;
    const size_t off = read_header(f, 0, c);
    long long bp = 0, naxis = 0;
    if (!off || !key_long(c, "BITPIX", &bp) || !key_long(c, "NAXIS", &naxis) ||
            naxis != 3)
        return false;
    for (int k = 0; k < 3; ++k)
        if (!key_long(c, "NAXIS" + std::to_string(k + 1), &dims[k]))
            return false;
    *bitpix = (int)bp;
    *data_off = off;
    return true;
}

void write_pixels(int precision, const void* data, const std::string& path,
                  int width, int height, int num_planes, int i_plane,
                  int* status)
{
    if (*status) return;
    const int bitpix = precision == GW_OSKAR_DOUBLE ? -64 : -32;
    int bp = 0;
    long long dims[3] = {0, 0, 0};
    size_t off = 0;
    // An existing cube of the same dimensions is updated in place (in its
    // own floating-point type); anything else is replaced.
    const bool dims_ok = image_geometry(path, &bp, dims, &off) &&
            (bp == -32 || bp == -64) && dims[0] == width &&
            dims[1] == height && dims[2] == num_planes;
    if (!dims_ok)
    {
        std::remove(path.c_str());
        if (!create_fits(path, bitpix, width, height, num_planes) ||
                !image_geometry(path, &bp, dims, &off))
        {
            *status = kErrFileIO;
            return;
        }
    }
    size_t first = (size_t)width * height * (size_t)std::max(i_plane, 0);
    size_t num = (size_t)width * height;
    if (i_plane < 0)
    {
        first = 0;
        num *= (size_t)num_planes;
    }
    if (!write_values(path, off, bp, first, num, precision, data))
        *status = kErrFileIO;
}

}  // namespace

extern "C" {

GW_IO_EXPORT gw_binary* gw_binary_open(const char* filename, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return nullptr;
    if (!filename)
    {
        *status = kErrInvalidArgument;
        return nullptr;
    }
    std::unique_ptr<gw_binary> h(new (std::nothrow) gw_binary());
    if (!h)
    {
        *status = kErrMemAlloc;
        return nullptr;
    }
    h->f = std::fopen(filename, "rb");
    if (!h->f)
    {
        *status = kBinOpenFail;
        return nullptr;
    }
    const int err = index_binary(*h);
    if (err)
    {
        std::fclose(h->f);
        *status = err;
        return nullptr;
    }
    return h.release();
}

GW_IO_EXPORT void gw_binary_free(gw_binary* h)
{
    if (!h) return;
    if (h->f) std::fclose(h->f);
    delete h;
}

GW_IO_EXPORT int gw_binary_num_tags(const gw_binary* h)
{
    return h ? (int)h->chunks.size() : 0;
}

GW_IO_EXPORT void gw_binary_set_query_search_start(gw_binary* h, int start,
        int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    if (!h)
    {
        *status = kErrInvalidArgument;
        return;
    }
    if (start < 0 || start >= (int)h->chunks.size())
    {
        *status = kBinTagOutOfRange;
        return;
    }
    h->search_start = start;
}

GW_IO_EXPORT int gw_binary_query(const gw_binary* h, unsigned char data_type,
        unsigned char id_group, unsigned char id_tag, int user_index,
        size_t* payload_size, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return 0;
    if (!h)
    {
        *status = kErrInvalidArgument;
        return -1;
    }
    for (int i = h->search_start; i < (int)h->chunks.size(); ++i)
    {
        const gw_binary::Chunk& c = h->chunks[i];
        if (!c.extended && (c.data_type == data_type || !data_type) &&
                c.id_group == id_group && c.id_tag == id_tag &&
                c.user_index == user_index)
        {
            if (payload_size) *payload_size = c.payload_size;
            return i;
        }
    }
    *status = kBinTagNotFound;
    return -1;
}

GW_IO_EXPORT void gw_binary_read(gw_binary* h, unsigned char data_type,
        unsigned char id_group, unsigned char id_tag, int user_index,
        size_t data_size, void* data, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    const int i = gw_binary_query(h, data_type, id_group, id_tag, user_index,
                                  nullptr, status);
    if (*status) return;
    read_chunk(h, i, data_size, data, status);
}

GW_IO_EXPORT void gw_binary_read_ext(gw_binary* h, unsigned char data_type,
        const char* name_group, const char* name_tag, int user_index,
        size_t data_size, void* data, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    if (!h || !name_group || !name_tag)
    {
        *status = kErrInvalidArgument;
        return;
    }
    if (std::strlen(name_group) + 1 > 255 || std::strlen(name_tag) + 1 > 255)
    {
        *status = kBinTagTooLong;
        return;
    }
    for (int i = h->search_start; i < (int)h->chunks.size(); ++i)
    {
        const gw_binary::Chunk& c = h->chunks[i];
        if (c.extended && (c.data_type == data_type || !data_type) &&
                c.name_group == name_group && c.name_tag == name_tag &&
                c.user_index == user_index)
        {
            read_chunk(h, i, data_size, data, status);
            return;
        }
    }
    *status = kBinTagNotFound;
}

GW_IO_EXPORT void gw_binary_read_int(gw_binary* h, unsigned char id_group,
        unsigned char id_tag, int user_index, int* value, int* status)
{
    gw_binary_read(h, GW_OSKAR_INT, id_group, id_tag, user_index,
                   sizeof(int), value, status);
}

GW_IO_EXPORT void gw_binary_read_double(gw_binary* h, unsigned char id_group,
        unsigned char id_tag, int user_index, double* value, int* status)
{
    gw_binary_read(h, GW_OSKAR_DOUBLE, id_group, id_tag, user_index,
                   sizeof(double), value, status);
}

GW_IO_EXPORT void gw_vis_read_coords(gw_binary* h, int i_block,
        int precision, size_t num_coords, void* uu, void* vv, void* ww,
        int* status)
{
    const size_t bytes = (precision == GW_OSKAR_DOUBLE ? sizeof(double) :
            sizeof(float)) * num_coords;
    gw_binary_read(h, (unsigned char)precision, GW_VIS_GROUP_BLOCK,
                   GW_VIS_BLOCK_TAG_BASELINE_UU, i_block, bytes, uu, status);
    gw_binary_read(h, (unsigned char)precision, GW_VIS_GROUP_BLOCK,
                   GW_VIS_BLOCK_TAG_BASELINE_VV, i_block, bytes, vv, status);
    gw_binary_read(h, (unsigned char)precision, GW_VIS_GROUP_BLOCK,
                   GW_VIS_BLOCK_TAG_BASELINE_WW, i_block, bytes, ww, status);
}

GW_IO_EXPORT void gw_vis_scale_coords(double freq_hz, int precision,
        size_t num_coords, void* uu, void* vv, void* ww)
{
    // Metres to wavelengths: the scale in FP64, applied in the precision of
    // the coordinates (read_vis.c:49-73).
    const double scale = freq_hz / 299792458.0;
    if (precision == GW_OSKAR_DOUBLE)
    {
        double* u = static_cast<double*>(uu);
        double* v = static_cast<double*>(vv);
        double* w = static_cast<double*>(ww);
        for (size_t i = 0; i < num_coords; ++i)
        {
            u[i] *= scale;
            v[i] *= scale;
            w[i] *= scale;
        }
    }
    else if (precision == GW_OSKAR_SINGLE)
    {
        const float s = (float)scale;
        float* u = static_cast<float*>(uu);
        float* v = static_cast<float*>(vv);
        float* w = static_cast<float*>(ww);
        for (size_t i = 0; i < num_coords; ++i)
        {
            u[i] *= s;
            v[i] *= s;
            w[i] *= s;
        }
    }
}

GW_IO_EXPORT void* gw_read_kernel(const char* file_name, int precision,
        int* conv_size_half, int* num_w_planes, int** support,
        int* oversample, int* grid_size, double* cellsize_rad,
        double* w_scale, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return nullptr;
    if (!file_name || !conv_size_half || !num_w_planes || !oversample ||
            !grid_size || !cellsize_rad || !w_scale)
    {
        *status = kErrInvalidArgument;
        return nullptr;
    }
    std::vector<unsigned char> f;
    if (!read_file(file_name, f))
    {
        *status = kErrFileIO;
        return nullptr;
    }
    std::vector<FitsCard> c;
    const size_t data_off = read_header(f, 0, c);
    long long bitpix = 0, naxis = 0, dims[3] = {0, 0, 0};
    if (!data_off || !key_long(c, "BITPIX", &bitpix) ||
            !key_long(c, "NAXIS", &naxis) || naxis != 3 ||
            !key_long(c, "NAXIS1", &dims[0]) ||
            !key_long(c, "NAXIS2", &dims[1]) ||
            !key_long(c, "NAXIS3", &dims[2]) || dims[0] < 1 || dims[1] < 1 ||
            dims[2] < 1)
    {
        *status = kErrFileIO;   // not a 3-D kernel cube
        return nullptr;
    }
    long long ov = 0, gs = 0;
    double cell = 0.0, ws = 0.0;
    if (!key_long(c, "OVERSAMP", &ov) || !key_long(c, "GRIDSIZE", &gs) ||
            !key_double(c, "CELLSIZE", &cell) ||
            !key_double(c, "W_SCALE", &ws))
    {
        *status = kErrFileIO;   // a header keyword is missing
        return nullptr;
    }
    const int bp = (int)bitpix;
    if (bp != 8 && bp != 16 && bp != 32 && bp != 64 && bp != -32 && bp != -64)
    {
        *status = kErrFileIO;
        return nullptr;
    }
    double bscale = 1.0, bzero = 0.0;
    key_double(c, "BSCALE", &bscale);
    key_double(c, "BZERO", &bzero);
    const size_t es = (size_t)std::abs(bp) / 8;
    size_t n = 0, nbytes = 0;
    if (__builtin_mul_overflow((size_t)dims[0], (size_t)dims[1], &n) ||
            __builtin_mul_overflow(n, (size_t)dims[2], &n) ||
            __builtin_mul_overflow(n, es, &nbytes) ||
            data_off > f.size() || nbytes > f.size() - data_off)
    {
        *status = kErrFileIO;   // the image does not fit the file
        return nullptr;
    }
    const size_t vs = precision == GW_OSKAR_DOUBLE ? 8 : 4;
    void* data = std::calloc(n, vs);
    if (!data)
    {
        *status = kErrMemAlloc;
        return nullptr;
    }
    const unsigned char* p = f.data() + data_off;
    for (size_t i = 0; i < n; ++i)
    {
        double x = be_value(p, bp, i);
        if (bscale != 1.0 || bzero != 0.0) x = x * bscale + bzero;
        if (vs == 8) static_cast<double*>(data)[i] = x;
        else static_cast<float*>(data)[i] = (float)x;
    }
    *conv_size_half = (int)dims[0];
    *num_w_planes = (int)dims[2];
    *oversample = (int)ov;
    *grid_size = (int)gs;
    *cellsize_rad = cell;
    *w_scale = ws;
    if (support)
    {
        // Column 1 of the W_KERNELS binary table, rows 1..num_w_planes.
        size_t pos = data_off + ((n * es + kBlock - 1) / kBlock) * kBlock;
        bool found = false;
        while (pos < f.size())
        {
            std::vector<FitsCard> e;
            const size_t doff = read_header(f, pos, e);
            if (!doff) break;
            long long nx1 = 0, nx2 = 0, pc = 0, gc = 1, bp2 = 8;
            key_long(e, "NAXIS1", &nx1);
            key_long(e, "NAXIS2", &nx2);
            key_long(e, "PCOUNT", &pc);
            key_long(e, "GCOUNT", &gc);
            key_long(e, "BITPIX", &bp2);
            const std::string* xt = find_key(e, "XTENSION");
            const std::string* name = find_key(e, "EXTNAME");
            const size_t dbytes = (size_t)(std::abs(bp2) / 8) *
                    (size_t)nx1 * (size_t)nx2 * (size_t)gc + (size_t)pc;
            if (xt && *xt == "BINTABLE" && name && *name == "W_KERNELS")
            {
                const std::string* tf = find_key(e, "TFORM1");
                if (!tf)
                    break;
                // Repeat count and type letter of column 1.
                size_t k = 0;
                while (k < tf->size() && std::isdigit((unsigned char)(*tf)[k]))
                    ++k;
                const char t = k < tf->size() ? (*tf)[k] : 'J';
                int cb = 0, cbp = 32;
                switch (t)
                {
                case 'B': cb = 1; cbp = 8; break;
                case 'I': cb = 2; cbp = 16; break;
                case 'J': cb = 4; cbp = 32; break;
                case 'K': cb = 8; cbp = 64; break;
                case 'E': cb = 4; cbp = -32; break;
                case 'D': cb = 8; cbp = -64; break;
                default: break;
                }
                double tscal = 1.0, tzero = 0.0;
                key_double(e, "TSCAL1", &tscal);
                key_double(e, "TZERO1", &tzero);
                // Rows of nx1 bytes, column 1 at the start of each row.
                size_t last = 0;
                if (!cb || nx1 < cb || nx2 < *num_w_planes ||
                        __builtin_mul_overflow((size_t)(*num_w_planes - 1),
                                               (size_t)nx1, &last) ||
                        doff > f.size() || last > f.size() - doff ||
                        (size_t)cb > f.size() - doff - last)
                    break;
                int* s = static_cast<int*>(std::calloc(*num_w_planes,
                                                       sizeof(int)));
                if (!s)
                {
                    std::free(data);
                    *status = kErrMemAlloc;
                    return nullptr;
                }
                // cfitsio's conversion to TINT without null checking
                // (getcolk.c): the scaled value, clamped to the int range
                // (an overflow fails the read), truncated toward zero; NaN
                // converts as x86 does (INT_MIN).
                bool overflow = false;
                for (int r = 0; r < *num_w_planes; ++r)
                {
                    const double x = be_value(f.data() + doff + (size_t)r * nx1,
                            cbp, 0) * tscal + tzero;
                    if (x < -2147483648.49) { overflow = true; s[r] = INT_MIN; }
                    else if (x > 2147483647.49) { overflow = true; s[r] = INT_MAX; }
                    else if (std::isnan(x)) s[r] = INT_MIN;
                    else s[r] = (int)x;
                }
                if (overflow)
                {
                    std::free(s);
                    break;
                }
                *support = s;
                found = true;
                break;
            }
            pos = doff + ((dbytes + kBlock - 1) / kBlock) * kBlock;
        }
        if (!found)
        {
            std::free(data);
            *status = kErrFileIO;   // the W_KERNELS table is missing
            return nullptr;
        }
    }
    return data;
}

GW_IO_EXPORT void gw_free(void* p)
{
    std::free(p);
}

GW_IO_EXPORT void gw_write_fits_cube(int data_type, const void* data,
        const char* root_name, int width, int height, int num_planes,
        int i_plane, int* status)
{
    int local = 0;
    if (!status) status = &local;
    if (*status) return;
    if (!data || !root_name || width < 1 || height < 1 || num_planes < 1 ||
            i_plane >= num_planes)
    {
        *status = kErrInvalidArgument;
        return;
    }
    const int precision = data_type & 0x0F;
    if (precision != GW_OSKAR_SINGLE && precision != GW_OSKAR_DOUBLE)
    {
        *status = kErrInvalidArgument;
        return;
    }
    const std::string root(root_name);
    if (data_type & GW_OSKAR_COMPLEX)
    {
        // Real and imaginary parts as two cubes, ROOT_REAL.fits and
        // ROOT_IMAG.fits (write_fits_cube.c:121-141).
        const size_t n = (size_t)width * height * num_planes;
        const size_t es = precision == GW_OSKAR_DOUBLE ? 8 : 4;
        std::vector<unsigned char> part(n * es);
        for (int c = 0; c < 2 && !*status; ++c)
        {
            const unsigned char* src = static_cast<const unsigned char*>(data);
            for (size_t i = 0; i < n; ++i)
                std::memcpy(&part[i * es], src + (2 * i + c) * es, es);
            write_pixels(precision, part.data(),
                         root + (c ? "_IMAG.fits" : "_REAL.fits"), width,
                         height, num_planes, i_plane, status);
        }
    }
    else
    {
        const size_t len = root.size();
        const bool has_ext = len >= 5 && (root.compare(len - 5, 5, ".fits") ==
                0 || root.compare(len - 5, 5, ".FITS") == 0);
        write_pixels(precision, data, has_ext ? root : root + ".fits", width,
                     height, num_planes, i_plane, status);
    }
}

}  // extern "C"
