// cc_io.cpp -- on-disk formats of the reference's CC tools (host only, no device): ConfigurationPack /
// SpherePackingPack files, the MSPB v1 snapshot, the MSPB archive and MultiStealthyPotential's text export.
//
//   ConfigurationPack   core/PeriodicCellList.h:1598-1640, core/PeriodicCellList.cpp:124-204
//                       Prefix.ConfigPackIndex: int64 count, then one int64 byte offset per record into Prefix.ConfigPack.
//                       A record is PeriodicCellList::WriteBinary (core/PeriodicCellList.h:1254-1267): uint16 dimension,
//                       double cell size, dimension x dimension doubles (basis vectors), int32 particle count, and per
//                       particle `dimension` doubles (relative coordinates) and its characteristic: char[4] (AtomName,
//                       core/PeriodicCellList.h:1558-1575) in a ConfigurationPack, one double (the radius) in a
//                       SpherePackingPack (core/SpherePacking.h:215-251, files Prefix.SpherePack / .SpherePackIndex).
//   MSPB v1             apps/CCOptimization/EXPORT_FORMAT.md (writer MultiStealthyPotential.cpp:46-81): little-endian;
//                       signature "MSPB\r\n\x1a\n", uint32 version 1, uint32 D, uint64 N, uint64 species count, uint32
//                       flags 3, D x D binary64 basis, N records (uint64 id, uint32 species, D relative coordinates,
//                       Cartesian radius).  Readers reject what EXPORT_FORMAT.md / MSPB_PACK.md list.
//   MSPB archive        apps/CCOptimization/MSPB_PACK.md: "MSPP\r\n\x1a\n", uint32 version 1, uint32 flags 0, then per
//                       snapshot a uint64 payload length and the unchanged MSPB bytes.
//   text export         MultiStealthyPotential.cpp:82-131: "# dimension", "# particles", "# basis_vector-d", coordinate
//                       and radius units, a tab-separated table (id, species, s_x.., radius), numbers at 17 significant
//                       digits (max_digits10, default float format).
// The pack files are written in the host's byte order as the reference does; every host here is little-endian.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <unordered_set>
#include <vector>

#include "cc_abi.h"

using namespace ccgpu;

namespace {

constexpr int kMaxDimension = 4;  // core/etc.h:12

template <class T>
void Put(std::string &b, T v)
{
    b.append(reinterpret_cast<const char *>(&v), sizeof(T));
}

template <class T>
bool Get(const std::string &b, size_t &off, T &v)
{
    if (off + sizeof(T) > b.size()) return false;
    std::memcpy(&v, b.data() + off, sizeof(T));
    off += sizeof(T);
    return true;
}

void PutLE(std::string &b, uint64_t v, int width)
{
    for (int i = 0; i < width; i++) {
        const unsigned char o = (unsigned char)(v & 0xff);
        b.push_back((char)o);
        v >>= 8;
    }
}

bool GetLE(const uint8_t *p, size_t size, size_t &off, int width, uint64_t &v)
{
    if (off + width > size) return false;
    v = 0;
    for (int i = 0; i < width; i++) v |= (uint64_t)p[off + i] << (8 * i);
    off += width;
    return true;
}

double Bits2Double(uint64_t u)
{
    double x;
    std::memcpy(&x, &u, 8);
    return x;
}

uint64_t Double2Bits(double x)
{
    uint64_t u;
    std::memcpy(&u, &x, 8);
    return u;
}

std::string ReadFile(const std::string &path, bool &ok)
{
    std::ifstream f(path, std::ios::binary);
    ok = f.good();
    if (!ok) return std::string();
    return std::string(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

void PackNames(const char *prefix, bool spheres, std::string &index, std::string &pack)
{
    const std::string p(prefix);
    index = p + (spheres ? ".SpherePackIndex" : ".ConfigPackIndex");
    pack = p + (spheres ? ".SpherePack" : ".ConfigPack");
}

// "%.17g" (an ostream at precision max_digits10 in the default float format)
void PutReal(std::string &out, double x)
{
    char buf[40];
    std::snprintf(buf, sizeof(buf), "%.17g", x);
    out += buf;
}

}  // namespace

extern "C" {

int ccgpu_pack_count(const char *prefix, int spheres, int64_t *count)
{
    return Guard([&] {
        Require(prefix && count, CCGPU_ERR_ARG, "null argument");
        std::string index, pack;
        PackNames(prefix, spheres != 0, index, pack);
        // ConfigurationPack::Open: a missing or empty index means no configuration
        std::ifstream f(index, std::ios::binary);
        long long n = 0;
        if (f.good()) {
            f.read(reinterpret_cast<char *>(&n), sizeof(n));
            if (!f) n = 0;
        }
        *count = n;
    });
}

int ccgpu_pack_add(const char *prefix, int spheres, int dim, double cell_size, const double *basis, int64_t N,
                   const double *relative, const void *chars)
{
    return Guard([&] {
        Require(prefix && basis && N >= 0 && (N == 0 || relative), CCGPU_ERR_ARG, "bad argument");
        Require(dim >= 1 && dim <= kMaxDimension, CCGPU_ERR_ARG, "dimension must be 1..4");
        Require(N <= 2147483647, CCGPU_ERR_ARG, "too many particles for the record's int32 count");
        std::string index, pack;
        PackNames(prefix, spheres != 0, index, pack);
        // the record (PeriodicCellList::WriteBinary)
        std::string rec;
        Put<uint16_t>(rec, (uint16_t)dim);
        Put<double>(rec, cell_size);
        for (int i = 0; i < dim * dim; i++) Put<double>(rec, basis[i]);
        Put<int32_t>(rec, (int32_t)N);
        const size_t cb = spheres ? sizeof(double) : 4;
        for (int64_t i = 0; i < N; i++) {
            for (int c = 0; c < dim; c++) Put<double>(rec, relative[i * dim + c]);
            if (chars) {
                rec.append(static_cast<const char *>(chars) + i * cb, cb);
            } else if (spheres) {
                Put<double>(rec, 0.0);
            } else {
                rec.append("a\0\0\0", 4);  // Insert("a", ...), as the reference's drivers name particles
            }
        }
        // ConfigurationPack::AddConfig: count at the head of the index, the record's offset appended to it
        int64_t n = 0;
        bool have = false;
        std::string idx = ReadFile(index, have);
        if (idx.size() >= 8) std::memcpy(&n, idx.data(), 8);
        else idx.assign(8, '\0');
        int64_t loc = 0;  // the record starts at the pack file's current end
        {
            std::ifstream pf(pack, std::ios::binary | std::ios::ate);
            if (pf.good()) loc = (int64_t)pf.tellg();
        }
        {
            std::ofstream pf(pack, std::ios::binary | std::ios::app);
            pf.write(rec.data(), (std::streamsize)rec.size());
            if (!pf) throw Error(CCGPU_ERR_STATE, "pack file write failed");
        }
        n++;
        std::memcpy(&idx[0], &n, 8);
        idx.append(reinterpret_cast<const char *>(&loc), 8);
        std::ofstream xf(index, std::ios::binary | std::ios::trunc);
        xf.write(idx.data(), (std::streamsize)idx.size());
        if (!xf) throw Error(CCGPU_ERR_STATE, "index file write failed");
    });
}

int ccgpu_pack_get(const char *prefix, int spheres, int64_t i, int *dim, double *cell_size, double *basis,
                   int64_t *N, int64_t cap, double *relative, void *chars)
{
    return Guard([&] {
        Require(prefix && dim && cell_size && basis && N, CCGPU_ERR_ARG, "null argument");
        std::string index, pack;
        PackNames(prefix, spheres != 0, index, pack);
        bool ok = false;
        const std::string idx = ReadFile(index, ok);
        int64_t n = 0;
        if (idx.size() >= 8) std::memcpy(&n, idx.data(), 8);
        if (i < 0 || i >= n) throw Error(CCGPU_ERR_RANGE, "configuration index out of range");
        int64_t loc = 0;
        size_t o = 8 * (size_t)(i + 1);
        if (!Get(idx, o, loc)) throw Error(CCGPU_ERR_STATE, "index file truncated");
        const std::string p = ReadFile(pack, ok);
        size_t off = (size_t)loc;
        uint16_t d16 = 0;
        double cs = 0;
        if (!Get(p, off, d16) || !Get(p, off, cs) || d16 > kMaxDimension || !(cs > 0.0))
            throw Error(CCGPU_ERR_STATE, "invalid configuration record header");
        const int d = d16;
        for (int k = 0; k < d * d; k++)
            if (!Get(p, off, basis[k])) throw Error(CCGPU_ERR_STATE, "configuration record truncated");
        int32_t np = 0;
        if (!Get(p, off, np)) throw Error(CCGPU_ERR_STATE, "configuration record truncated");
        if (np < 0) np = 0;  // a corrupt count reads as empty, as the reference's reader does
        *dim = d;
        *cell_size = cs;
        *N = np;
        const size_t cb = spheres ? sizeof(double) : 4;
        const int64_t w = std::min<int64_t>(np, std::max<int64_t>(cap, 0));
        for (int64_t q = 0; q < np; q++) {
            for (int c = 0; c < d; c++) {
                double v;
                if (!Get(p, off, v)) throw Error(CCGPU_ERR_STATE, "configuration record truncated");
                if (q < w && relative) relative[q * d + c] = v;
            }
            if (off + cb > p.size()) throw Error(CCGPU_ERR_STATE, "configuration record truncated");
            if (q < w && chars) std::memcpy(static_cast<char *>(chars) + q * cb, p.data() + off, cb);
            off += cb;
        }
    });
}

int ccgpu_pack_clear(const char *prefix, int spheres)
{
    return Guard([&] {
        Require(prefix != nullptr, CCGPU_ERR_ARG, "null argument");
        std::string index, pack;
        PackNames(prefix, spheres != 0, index, pack);
        std::ofstream a(index, std::ios::binary | std::ios::trunc), b(pack, std::ios::binary | std::ios::trunc);
    });
}

int ccgpu_mspb_encode(int dim, const double *basis, int64_t N, const uint64_t *ids, const uint32_t *species,
                      const double *relative, const double *radii, uint64_t nspecies, uint8_t *buf, int64_t cap,
                      int64_t *size)
{
    return Guard([&] {
        Require(basis && size && N >= 0 && (N == 0 || (species && relative && radii)), CCGPU_ERR_ARG, "bad argument");
        Require(dim >= 1 && dim <= kMaxDimension, CCGPU_ERR_ARG, "dimension must be 1..4");
        std::string b("MSPB\r\n\x1a\n", 8);
        auto real = [&](double x) {
            if (!std::isfinite(x)) throw Error(CCGPU_ERR_ARG, "nonfinite binary value");
            PutLE(b, Double2Bits(x), 8);
        };
        PutLE(b, 1, 4);
        PutLE(b, (uint64_t)dim, 4);
        PutLE(b, (uint64_t)N, 8);
        PutLE(b, nspecies, 8);
        PutLE(b, 3, 4);
        for (int k = 0; k < dim * dim; k++) real(basis[k]);
        for (int64_t i = 0; i < N; i++) {
            Require(species[i] < nspecies, CCGPU_ERR_ARG, "species label out of range");
            Require(std::isfinite(radii[i]) && radii[i] > 0.0, CCGPU_ERR_ARG, "invalid radius");
            PutLE(b, ids ? ids[i] : (uint64_t)i, 8);
            PutLE(b, species[i], 4);
            for (int c = 0; c < dim; c++) real(relative[i * dim + c]);
            real(radii[i]);
        }
        *size = (int64_t)b.size();
        if (buf && cap >= (int64_t)b.size()) std::memcpy(buf, b.data(), b.size());
    });
}

int ccgpu_mspb_decode(const uint8_t *p, int64_t size, int *dim, double *basis, int64_t *N, uint64_t *nspecies,
                      int64_t cap, uint64_t *ids, uint32_t *species, double *relative, double *radii)
{
    return Guard([&] {
        Require(p && size >= 0 && dim && basis && N && nspecies, CCGPU_ERR_ARG, "null argument");
        const size_t sz = (size_t)size;
        size_t off = 0;
        if (sz < 8 || std::memcmp(p, "MSPB\r\n\x1a\n", 8) != 0) throw Error(CCGPU_ERR_ARG, "not an MSPB snapshot");
        off = 8;
        uint64_t ver, d, n, ns, flags;
        if (!GetLE(p, sz, off, 4, ver) || !GetLE(p, sz, off, 4, d) || !GetLE(p, sz, off, 8, n) ||
            !GetLE(p, sz, off, 8, ns) || !GetLE(p, sz, off, 4, flags))
            throw Error(CCGPU_ERR_ARG, "MSPB header truncated");
        if (ver != 1) throw Error(CCGPU_ERR_ARG, "unsupported MSPB version");
        if (flags != 3) throw Error(CCGPU_ERR_ARG, "unsupported MSPB flags");
        if (d < 1 || d > (uint64_t)kMaxDimension) throw Error(CCGPU_ERR_ARG, "MSPB dimension out of range");
        const uint64_t rec = 20 + 8 * d;
        if (n > (sz - std::min<size_t>(sz, 36 + 8 * d * d)) / rec || 36 + 8 * d * d + n * rec != sz)
            throw Error(CCGPU_ERR_ARG, "MSPB byte count does not match its header (truncated or trailing data)");
        *dim = (int)d;
        *N = (int64_t)n;
        *nspecies = ns;
        double a[16];
        for (uint64_t k = 0; k < d * d; k++) {
            uint64_t u;
            GetLE(p, sz, off, 8, u);
            a[k] = Bits2Double(u);
            if (!std::isfinite(a[k])) throw Error(CCGPU_ERR_ARG, "nonfinite MSPB basis");
            basis[k] = a[k];
        }
        const double vol = CellVolume((int)d, a);
        if (!(vol > 0.0) || !std::isfinite(vol)) throw Error(CCGPU_ERR_ARG, "singular MSPB cell");
        std::unordered_set<uint64_t> seen;
        const int64_t w = std::min<int64_t>((int64_t)n, std::max<int64_t>(cap, 0));
        for (uint64_t i = 0; i < n; i++) {
            uint64_t id, lab, u;
            GetLE(p, sz, off, 8, id);
            GetLE(p, sz, off, 4, lab);
            if (!seen.insert(id).second) throw Error(CCGPU_ERR_ARG, "duplicate MSPB particle id");
            if (lab >= ns) throw Error(CCGPU_ERR_ARG, "MSPB species label out of range");
            if ((int64_t)i < w) {
                if (ids) ids[i] = id;
                if (species) species[i] = (uint32_t)lab;
            }
            for (uint64_t c = 0; c < d; c++) {
                GetLE(p, sz, off, 8, u);
                const double x = Bits2Double(u);
                if (!std::isfinite(x)) throw Error(CCGPU_ERR_ARG, "nonfinite MSPB coordinate");
                if ((int64_t)i < w && relative) relative[i * d + c] = x;
            }
            GetLE(p, sz, off, 8, u);
            const double r = Bits2Double(u);
            if (!std::isfinite(r) || !(r > 0.0)) throw Error(CCGPU_ERR_ARG, "invalid MSPB radius");
            if ((int64_t)i < w && radii) radii[i] = r;
        }
    });
}

int ccgpu_export_text(int dim, const double *basis, int64_t N, const int32_t *species, const double *relative,
                      const double *radii, char *buf, int64_t cap, int64_t *size)
{
    return Guard([&] {
        Require(basis && size && N >= 0 && (N == 0 || (species && relative && radii)), CCGPU_ERR_ARG, "bad argument");
        Require(dim >= 1 && dim <= kMaxDimension, CCGPU_ERR_ARG, "dimension must be 1..4");
        std::string t;
        t += "# dimension " + std::to_string(dim) + "\n";
        t += "# particles " + std::to_string(N) + "\n";
        for (int q = 0; q < dim; q++) {
            t += "# basis_vector-" + std::to_string(q);
            for (int c = 0; c < dim; c++) {
                if (!std::isfinite(basis[q * dim + c])) throw Error(CCGPU_ERR_ARG, "nonfinite basis vector");
                t += ' ';
                PutReal(t, basis[q * dim + c]);
            }
            t += '\n';
        }
        t += "# coordinates relative\n# radius_units cartesian\n";
        t += "# id\tspecies";
        for (int c = 0; c < dim; c++) {
            if (c < 3) {
                t += "\ts";
                t += "xyz"[c];
            } else {
                t += "\ts" + std::to_string(c);
            }
        }
        t += "\tradius\n";
        for (int64_t i = 0; i < N; i++) {
            Require(species[i] >= 0, CCGPU_ERR_ARG, "particle has no valid species");
            Require(std::isfinite(radii[i]) && radii[i] > 0.0, CCGPU_ERR_ARG, "particle has an invalid radius");
            t += std::to_string(i) + "\t" + std::to_string(species[i]);
            for (int c = 0; c < dim; c++) {
                if (!std::isfinite(relative[i * dim + c])) throw Error(CCGPU_ERR_ARG, "nonfinite coordinate");
                t += '\t';
                PutReal(t, relative[i * dim + c]);
            }
            t += '\t';
            PutReal(t, radii[i]);
            t += '\n';
        }
        *size = (int64_t)t.size();
        if (buf && cap >= (int64_t)t.size()) std::memcpy(buf, t.data(), t.size());
    });
}

int ccgpu_mspbpack_count(const char *prefix, int64_t *count)
{
    return Guard([&] {
        Require(prefix && count, CCGPU_ERR_ARG, "null argument");
        bool ok = false;
        const std::string a = ReadFile(std::string(prefix) + ".MSPBPack", ok);
        *count = 0;
        if (!ok) return;  // missing archive: empty
        size_t off = 0;
        const uint8_t *p = reinterpret_cast<const uint8_t *>(a.data());
        uint64_t ver, flags;
        if (a.size() < 16 || std::memcmp(p, "MSPP\r\n\x1a\n", 8) != 0) throw Error(CCGPU_ERR_STATE, "not an MSPB archive");
        off = 8;
        GetLE(p, a.size(), off, 4, ver);
        GetLE(p, a.size(), off, 4, flags);
        if (ver != 1 || flags != 0) throw Error(CCGPU_ERR_STATE, "unsupported MSPB archive version or flags");
        int64_t n = 0;
        while (off < a.size()) {
            uint64_t len;
            if (!GetLE(p, a.size(), off, 8, len) || len > a.size() - off)
                throw Error(CCGPU_ERR_STATE, "MSPB archive truncated");
            off += len;
            n++;
        }
        *count = n;
    });
}

int ccgpu_mspbpack_add(const char *prefix, const uint8_t *snapshot, int64_t size)
{
    return Guard([&] {
        Require(prefix && snapshot && size > 0, CCGPU_ERR_ARG, "bad argument");
        // validate the payload as a reader would before appending it
        int d;
        double basis[16];
        int64_t n;
        uint64_t ns;
        const int st = ccgpu_mspb_decode(snapshot, size, &d, basis, &n, &ns, 0, nullptr, nullptr, nullptr, nullptr);
        if (st != CCGPU_OK) throw Error(st, LastError());
        int64_t count = 0;
        const int cs = ccgpu_mspbpack_count(prefix, &count);
        if (cs != CCGPU_OK) throw Error(cs, LastError());
        const std::string path = std::string(prefix) + ".MSPBPack";
        bool ok = false;
        const std::string existing = ReadFile(path, ok);
        std::string out;
        if (!ok || existing.empty()) {
            out.assign("MSPP\r\n\x1a\n", 8);
            PutLE(out, 1, 4);
            PutLE(out, 0, 4);
        }
        PutLE(out, (uint64_t)size, 8);
        out.append(reinterpret_cast<const char *>(snapshot), (size_t)size);
        std::ofstream f(path, std::ios::binary | std::ios::app);
        f.write(out.data(), (std::streamsize)out.size());
        if (!f) throw Error(CCGPU_ERR_STATE, "MSPB archive write failed");
    });
}

int ccgpu_mspbpack_get(const char *prefix, int64_t index, uint8_t *buf, int64_t cap, int64_t *size)
{
    return Guard([&] {
        Require(prefix && size, CCGPU_ERR_ARG, "null argument");
        int64_t count = 0;
        const int cs = ccgpu_mspbpack_count(prefix, &count);
        if (cs != CCGPU_OK) throw Error(cs, LastError());
        if (index < 0 || index >= count) throw Error(CCGPU_ERR_RANGE, "snapshot index out of range");
        bool ok = false;
        const std::string a = ReadFile(std::string(prefix) + ".MSPBPack", ok);
        const uint8_t *p = reinterpret_cast<const uint8_t *>(a.data());
        size_t off = 16;
        for (int64_t i = 0;; i++) {
            uint64_t len;
            GetLE(p, a.size(), off, 8, len);
            if (i == index) {
                *size = (int64_t)len;
                if (buf && cap >= (int64_t)len) std::memcpy(buf, p + off, len);
                return;
            }
            off += len;
        }
    });
}

}  // extern "C"

// MD restart file (CollectiveCoordinateMD, apps/CCOptimization/main.cpp:1891-1908, 1936-1946, 1973-1983, 2005-2011;
// ParticleMolecularDynamics::WriteBinary and its stream constructor, potential/MD_System.cpp:34-63): int32 stage,
// int64 step, the Position record (PeriodicCellList::WriteBinary, as in a ConfigurationPack), double TimeStep,
// double Time, uint64 NumParticle (size_t), uint16 Dimension, double xi, double MassQ, NumParticle doubles Mass, and
// NumParticle GeometryVector for Velocity and for Acceleration, each the four doubles x[MaxDimension] of a release
// build (core/GeometryVector.h:24-47: the recorded dimension exists only without NDEBUG), components beyond the
// dimension zero.  Velocity and acceleration cross the ABI as N x dim doubles.
int ccgpu_mddump_write(const char *path, int32_t stage, int64_t step, int dim, double cell_size, const double *basis,
                       int64_t N, const double *relative, const void *chars, double timestep, double time, double xi,
                       double Q, const double *mass, const double *velocity, const double *acceleration)
{
    return Guard([&] {
        Require(path && basis && N >= 0 && (N == 0 || (relative && mass && velocity && acceleration)), CCGPU_ERR_ARG,
                "bad argument");
        Require(dim >= 1 && dim <= kMaxDimension, CCGPU_ERR_ARG, "dimension must be 1..4");
        Require(N <= 2147483647, CCGPU_ERR_ARG, "too many particles for the record's int32 count");
        std::string b;
        Put<int32_t>(b, stage);
        Put<int64_t>(b, step);
        Put<uint16_t>(b, (uint16_t)dim);
        Put<double>(b, cell_size);
        for (int i = 0; i < dim * dim; i++) Put<double>(b, basis[i]);
        Put<int32_t>(b, (int32_t)N);
        for (int64_t i = 0; i < N; i++) {
            for (int c = 0; c < dim; c++) Put<double>(b, relative[i * dim + c]);
            if (chars) b.append(static_cast<const char *>(chars) + i * 4, 4);
            else b.append("a\0\0\0", 4);
        }
        Put<double>(b, timestep);
        Put<double>(b, time);
        Put<uint64_t>(b, (uint64_t)N);
        Put<uint16_t>(b, (uint16_t)dim);
        Put<double>(b, xi);
        Put<double>(b, Q);
        for (int64_t i = 0; i < N; i++) Put<double>(b, mass[i]);
        for (const double *vec : {velocity, acceleration})
            for (int64_t i = 0; i < N; i++)
                for (int c = 0; c < kMaxDimension; c++) Put<double>(b, c < dim ? vec[i * dim + c] : 0.0);
        std::ofstream f(path, std::ios::binary | std::ios::trunc);
        f.write(b.data(), (std::streamsize)b.size());
        if (!f) throw Error(CCGPU_ERR_STATE, "MDDump write failed");
    });
}

int ccgpu_mddump_read(const char *path, int32_t *stage, int64_t *step, int *dim, double *cell_size, double *basis,
                      int64_t *N, int64_t cap, double *relative, void *chars, double *timestep, double *time,
                      double *xi, double *Q, double *mass, double *velocity, double *acceleration)
{
    return Guard([&] {
        Require(path && stage && step && dim && cell_size && basis && N && timestep && time && xi && Q, CCGPU_ERR_ARG,
                "null argument");
        bool ok = false;
        const std::string p = ReadFile(path, ok);
        if (!ok) throw Error(CCGPU_ERR_STATE, std::string("cannot read ") + path);
        size_t off = 0;
        int32_t st = 0;
        int64_t sp = 0;
        uint16_t d16 = 0;
        double cs = 0;
        if (!Get(p, off, st) || !Get(p, off, sp) || !Get(p, off, d16) || !Get(p, off, cs) || d16 < 1 ||
            d16 > kMaxDimension)
            throw Error(CCGPU_ERR_STATE, "invalid MDDump header");
        const int d = d16;
        for (int k = 0; k < d * d; k++)
            if (!Get(p, off, basis[k])) throw Error(CCGPU_ERR_STATE, "MDDump truncated");
        int32_t np = 0;
        if (!Get(p, off, np) || np < 0) throw Error(CCGPU_ERR_STATE, "invalid MDDump particle count");
        *stage = st;
        *step = sp;
        *dim = d;
        *cell_size = cs;
        *N = np;
        const bool fill = cap >= np;  // the arrays are filled only when they hold every particle
        for (int64_t q = 0; q < np; q++) {
            for (int c = 0; c < d; c++) {
                double v;
                if (!Get(p, off, v)) throw Error(CCGPU_ERR_STATE, "MDDump truncated");
                if (fill && relative) relative[q * d + c] = v;
            }
            if (off + 4 > p.size()) throw Error(CCGPU_ERR_STATE, "MDDump truncated");
            if (fill && chars) std::memcpy(static_cast<char *>(chars) + q * 4, p.data() + off, 4);
            off += 4;
        }
        uint64_t n2 = 0;
        uint16_t d2 = 0;
        if (!Get(p, off, *timestep) || !Get(p, off, *time) || !Get(p, off, n2) || !Get(p, off, d2) ||
            !Get(p, off, *xi) || !Get(p, off, *Q))
            throw Error(CCGPU_ERR_STATE, "MDDump truncated");
        if (n2 != (uint64_t)np || d2 != d16) throw Error(CCGPU_ERR_STATE, "MDDump particle count or dimension mismatch");
        for (int64_t q = 0; q < np; q++) {
            double m;
            if (!Get(p, off, m)) throw Error(CCGPU_ERR_STATE, "MDDump truncated");
            if (fill && mass) mass[q] = m;
        }
        for (double *vec : {velocity, acceleration})
            for (int64_t q = 0; q < np; q++)
                for (int c = 0; c < kMaxDimension; c++) {
                    double v;
                    if (!Get(p, off, v)) throw Error(CCGPU_ERR_STATE, "MDDump truncated");
                    if (fill && vec && c < d) vec[q * d + c] = v;
                }
        if (off != p.size()) throw Error(CCGPU_ERR_STATE, "MDDump has trailing bytes");
    });
}
