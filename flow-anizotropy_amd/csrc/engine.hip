// engine.hip -- geometry upload, job lattices and the chunked batch stepper.
#include "engine.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>

namespace fp {

double now_seconds()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static std::string hip_err(const char *what, hipError_t e)
{
    return std::string(what) + ": " + hipGetErrorString(e);
}

Geometry::~Geometry()
{
    if(d_seg) (void)hipFree(d_seg);
    if(d_links) (void)hipFree(d_links);
    if(d_rowspec) (void)hipFree(d_rowspec);
    if(d_ctab) (void)hipFree(d_ctab);
    if(d_rowbase) (void)hipFree(d_rowbase);
}

size_t Geometry::compact_bytes() const
{
    const size_t rec = dim == 3 ? sizeof(CRec<RowSlots<D3Q19>::N>) : sizeof(CRec<RowSlots<D2Q9>::N>);
    const int words = dim == 3 ? rowtab_words<D3Q19>() : rowtab_words<D2Q9>();
    return (size_t)n_fluid * rec + (size_t)nrows * words * sizeof(uint32_t);
}

static int64_t round_up(int64_t v, int64_t m);

// The compact neighbour table (CRec per fluid cell, start and last rank of
// every row slot per row) from the segment table.  False when a row holds
// 0xFFFF or more cells (the ranks are 16-bit); the row-segment neighbour step
// is used then.
template<class L>
static bool build_compact(const Geometry &g, const std::vector<SegInfo> &h,
                          std::vector<CRec<RowSlots<L>::N>> &tab, std::vector<uint32_t> &rowtab)
{
    constexpr int NR = RowSlots<L>::N;
    const int NW = rowtab_words<L>();
    if(g.nx >= 0xFFFF) return false;
    auto wrap = [](int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); };
    auto rowstart = [&](int64_t r) -> uint32_t { return h[(size_t)(r * g.nseg)].start; };
    auto rowcount = [&](int64_t r) -> uint32_t {
        return (uint32_t)((r + 1 < g.nrows ? (int64_t)rowstart(r + 1) : g.n_fluid) - rowstart(r));
    };
    tab.assign((size_t)g.n_fluid, CRec<NR>{});
    rowtab.assign((size_t)g.nrows * NW, 0u);
    for(int64_t row = 0; row < g.nrows; row++) {
        const int y = L::DIM == 3 ? (int)(row % g.ny) : (int)row;
        const int z = L::DIM == 3 ? (int)(row / g.ny) : 0;
        const uint8_t *rs[NR];
        for(int m = 0; m < NR; m++) {
            const int yy = wrap(y - RowSlots<L>::ey(m), g.ny);
            const int64_t r = L::DIM == 3 ? (int64_t)wrap(z - RowSlots<L>::ez(m), g.nz) * g.ny + yy : yy;
            rs[m] = g.solid.data() + r * g.nx;
            const uint32_t cnt = rowcount(r);
            rowtab[(size_t)row * NW + m] = rowstart(r);
            rowtab[(size_t)row * NW + NR + m] = cnt ? cnt - 1u : 0u;
        }
        uint32_t P[NR] = {};                  // fluid cells of each slot's row before x
        uint32_t i = rowstart(row);
        for(int x = 0; x < g.nx; x++) {
            if(!rs[0][x]) {
                CRec<NR> &R = tab[i++];
                R.row = (uint32_t)row;
                uint32_t flags = 0;
                for(int m = 0; m < NR; m++)
                    for(int e = 0; e < 3; e++)
                        if(!rs[m][wrap(x - 1 + e, g.nx)]) flags |= 1u << (3 * m + e);
                if(x == 0) flags |= CFLAG_X0;
                if(x == g.nx - 1) flags |= CFLAG_XLAST;
                R.flags = flags;
                for(int m = 1; m < NR; m++) R.rank[m - 1] = (uint16_t)P[m];
            }
            for(int m = 0; m < NR; m++) P[m] += rs[m][x] == 0;
        }
    }
    return true;
}

// The neighbour links of every segment (LinkRec): which lanes' neighbours
// along each direction are fluid and where they sit in the compact order.
template<class L>
static void build_links(const Geometry &g, const std::vector<SegInfo> &h,
                        std::vector<LinkRec<L::Q - 1>> &links, std::vector<uint32_t> &rowspec)
{
    constexpr int NK = L::Q - 1;
    const int nseg = g.nseg, lw = (g.nx - 1) & 63;
    links.resize((size_t)g.nunits);
    rowspec.assign((size_t)g.nrows * NK, 0u);
    auto wrap = [](int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); };
    for(int64_t row = 0; row < g.nrows; row++) {
        const int y = L::DIM == 3 ? (int)(row % g.ny) : (int)row;
        const int z = L::DIM == 3 ? (int)(row / g.ny) : 0;
        for(int sg = 0; sg < nseg; sg++) {
            LinkRec<NK> &R = links[(size_t)(row * nseg + sg)];
            for(int k = 1; k < L::Q; k++) {
                const int dx = -L::ex(k), dy = -L::ey(k), dz = -L::ez(k);
                const int yy = wrap(y + dy, g.ny);
                const int64_t rr = L::DIM == 3 ? (int64_t)wrap(z + dz, g.nz) * g.ny + yy : yy;
                const SegInfo &c = h[(size_t)(rr * nseg + sg)];
                uint64_t M;
                uint32_t base;
                if(dx == 0) {
                    M = c.mask;
                    base = c.start;
                } else if(dx < 0) {
                    // x - 1: this segment shifted up one lane; lane 0 takes the
                    // previous segment's last cell (the row's x = nx - 1 at sg = 0)
                    const int sp = sg > 0 ? sg - 1 : nseg - 1;
                    const SegInfo &pv = h[(size_t)(rr * nseg + sp)];
                    const uint64_t prevbit = sg > 0 ? (pv.mask >> 63) & 1ull : (pv.mask >> lw) & 1ull;
                    M = (c.mask << 1) | prevbit;
                    base = c.start - (uint32_t)prevbit;
                    if(sg == 0)
                        rowspec[(size_t)row * NK + k - 1] =
                            pv.start + (uint32_t)__builtin_popcountll(pv.mask & ((1ull << lw) - 1ull));
                } else {
                    // x + 1: shifted down one lane; the segment's last lane takes
                    // the next segment's first cell (the row's x = 0 at the end)
                    const int ls = sg < nseg - 1 ? 63 : lw;
                    const int sn = sg < nseg - 1 ? sg + 1 : 0;
                    const SegInfo &nx_ = h[(size_t)(rr * nseg + sn)];
                    M = ((c.mask >> 1) & ~(1ull << ls)) | ((nx_.mask & 1ull) << ls);
                    base = c.start + (uint32_t)(c.mask & 1ull);
                    if(sg == nseg - 1) rowspec[(size_t)row * NK + k - 1] = nx_.start;
                }
                R.M[k - 1] = M;
                R.base[k - 1] = base;
            }
        }
    }
}

template<class L>
static std::string upload_links(Geometry &g, const std::vector<SegInfo> &h)
{
    std::vector<LinkRec<L::Q - 1>> links;
    std::vector<uint32_t> rowspec;
    build_links<L>(g, h, links, rowspec);
    g.links_bytes = links.size() * sizeof(links[0]);
    g.rowspec_bytes = rowspec.size() * sizeof(uint32_t);
    hipError_t e = hipMalloc(&g.d_links, g.links_bytes);
    if(e != hipSuccess) { g.d_links = nullptr; return std::string("hipMalloc(links): ") + hipGetErrorString(e); }
    e = hipMalloc((void **)&g.d_rowspec, g.rowspec_bytes);
    if(e != hipSuccess) { g.d_rowspec = nullptr; return std::string("hipMalloc(links): ") + hipGetErrorString(e); }
    e = hipMemcpy(g.d_links, links.data(), g.links_bytes, hipMemcpyHostToDevice);
    if(e == hipSuccess) e = hipMemcpy(g.d_rowspec, rowspec.data(), g.rowspec_bytes, hipMemcpyHostToDevice);
    if(e != hipSuccess) return std::string("hipMemcpy(links): ") + hipGetErrorString(e);
    return "";
}

// host half of geometry_build: dimensions, n_fluid and the segment table
static std::string build_segments(Geometry &g, const uint8_t *solid, int dim, int nx, int ny, int nz,
                                  std::vector<SegInfo> &h)
{
    if(nx <= 0 || ny <= 0 || (dim == 3 && nz <= 0)) return "bad dimensions";
    g.dim = dim;
    g.nx = nx; g.ny = ny; g.nz = dim == 3 ? nz : 1;
    g.nseg = (nx + 63) / 64;
    g.nrows = (int64_t)g.ny * g.nz;
    g.nunits = g.nrows * g.nseg;
    g.nn = (int64_t)nx * g.nrows;
    if(g.nunits > INT32_MAX || g.nseg > INT16_MAX) return "lattice too large";
    h.assign((size_t)g.nunits, SegInfo{});
    uint64_t count = 0;
    for(int64_t row = 0; row < g.nrows; row++) {
        const uint8_t *r = solid + row * nx;
        for(int sg = 0; sg < g.nseg; sg++) {
            uint64_t m = 0;
            const int x0 = sg * 64, x1 = std::min(nx, x0 + 64);
            for(int x = x0; x < x1; x++) {
                if(r[x] > 1) return "non-binary voxel";
                if(r[x] == 0) m |= 1ull << (x - x0);
            }
            SegInfo &s = h[(size_t)(row * g.nseg + sg)];
            s.mask = m;
            s.start = (uint32_t)count;
            s.pad = 0;
            count += (uint64_t)__builtin_popcountll(m);
        }
    }
    if(count == 0) return "structure has no pore space";
    if(count >= (1ull << 32)) return "too many fluid cells for 32-bit indices";
    g.n_fluid = (int64_t)count;
    return "";
}

std::string geometry_host_tables(Geometry &g, const uint8_t *solid, int dim, int nx, int ny, int nz,
                                 std::vector<SegInfo> &seg, std::vector<unsigned char> &links,
                                 std::vector<uint32_t> &rowspec)
{
    std::string err = build_segments(g, solid, dim, nx, ny, nz, seg);
    if(!err.empty()) return err;
    if(dim == 3) {
        std::vector<LinkRec<D3Q19::Q - 1>> l;
        build_links<D3Q19>(g, seg, l, rowspec);
        links.assign((const unsigned char *)l.data(), (const unsigned char *)(l.data() + l.size()));
    } else {
        std::vector<LinkRec<D2Q9::Q - 1>> l;
        build_links<D2Q9>(g, seg, l, rowspec);
        links.assign((const unsigned char *)l.data(), (const unsigned char *)(l.data() + l.size()));
    }
    g.links_bytes = links.size();
    g.rowspec_bytes = rowspec.size() * sizeof(uint32_t);
    g.solid.assign(solid, solid + g.nn);
    g.row_first.resize((size_t)g.nrows + 1);
    for(int64_t r = 0; r < g.nrows; r++) g.row_first[(size_t)r] = seg[(size_t)(r * g.nseg)].start;
    g.row_first[(size_t)g.nrows] = (uint32_t)g.n_fluid;
    return "";
}

std::string geometry_build(Geometry &g, const uint8_t *solid, int dim, int nx, int ny, int nz)
{
    std::vector<SegInfo> h;
    std::string err = build_segments(g, solid, dim, nx, ny, nz, h);
    if(!err.empty()) return err;
    hipError_t e = hipMalloc(&g.d_seg, h.size() * sizeof(SegInfo));
    if(e != hipSuccess) { g.d_seg = nullptr; return hip_err("hipMalloc(segments)", e); }
    e = hipMemcpy(g.d_seg, h.data(), h.size() * sizeof(SegInfo), hipMemcpyHostToDevice);
    if(e != hipSuccess) return hip_err("hipMemcpy(segments)", e);
    std::string le = dim == 3 ? upload_links<D3Q19>(g, h) : upload_links<D2Q9>(g, h);
    if(!le.empty()) return le;
    g.solid.assign(solid, solid + g.nn);
    g.row_first.resize((size_t)g.nrows + 1);
    for(int64_t r = 0; r < g.nrows; r++) g.row_first[(size_t)r] = h[(size_t)(r * g.nseg)].start;
    g.row_first[(size_t)g.nrows] = (uint32_t)g.n_fluid;
    return "";
}

// The neighbour index the step kernel derives from the link tables, for
// every fluid cell and direction, against a direct lookup on the dense
// lattice (periodic neighbour, compact index by counting fluid cells).
template<class L>
static int64_t verify_links(const Geometry &g, const uint8_t *solid, const std::vector<SegInfo> &h)
{
    constexpr int NK = L::Q - 1;
    std::vector<LinkRec<NK>> links;
    std::vector<uint32_t> rowspec;
    build_links<L>(g, h, links, rowspec);
    std::vector<int64_t> cidx((size_t)g.nn, -1);
    int64_t c = 0;
    for(int64_t n = 0; n < g.nn; n++) if(solid[n] == 0) cidx[(size_t)n] = c++;
    int64_t bad = 0;
    const int lw = (g.nx - 1) & 63;
    for(int64_t row = 0; row < g.nrows; row++) {
        const int y = L::DIM == 3 ? (int)(row % g.ny) : (int)row;
        const int z = L::DIM == 3 ? (int)(row / g.ny) : 0;
        for(int sg = 0; sg < g.nseg; sg++) {
            const LinkRec<NK> &R = links[(size_t)(row * g.nseg + sg)];
            const SegInfo &own = h[(size_t)(row * g.nseg + sg)];
            for(int lane = 0; lane < 64; lane++) {
                if(!((own.mask >> lane) & 1ull)) continue;
                const int x = sg * 64 + lane;
                const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
                for(int k = 1; k < L::Q; k++) {
                    const uint64_t M = R.M[k - 1];
                    uint32_t nbi = R.base[k - 1] + (uint32_t)__builtin_popcountll(M & below);
                    if(L::ex(k) > 0 && sg == 0 && lane == 0) nbi = rowspec[(size_t)row * NK + k - 1];
                    if(L::ex(k) < 0 && sg == g.nseg - 1 && lane == lw) nbi = rowspec[(size_t)row * NK + k - 1];
                    const bool fluid = (M >> lane) & 1ull;
                    const int xx = (x - L::ex(k) + g.nx) % g.nx;
                    const int yy = (y - L::ey(k) + g.ny) % g.ny;
                    const int zz = L::DIM == 3 ? (z - L::ez(k) + g.nz) % g.nz : 0;
                    const int64_t dn = ((int64_t)zz * g.ny + yy) * g.nx + xx;
                    const bool want_fluid = solid[(size_t)dn] == 0;
                    if(fluid != want_fluid || (fluid && (int64_t)nbi != cidx[(size_t)dn])) bad++;
                }
            }
        }
    }
    return bad;
}

// the compact table against a dense lookup: the slot the compact neighbour
// step addresses for every fluid cell and direction
template<class L>
static int64_t verify_compact(const Geometry &g, const uint8_t *solid, const std::vector<SegInfo> &h)
{
    constexpr int NR = RowSlots<L>::N;
    const int NW = rowtab_words<L>();
    std::vector<CRec<NR>> tab;
    std::vector<uint32_t> rowtab;
    if(!build_compact<L>(g, h, tab, rowtab)) return 0;
    std::vector<int64_t> cidx((size_t)g.nn, -1);
    int64_t c = 0;
    for(int64_t n = 0; n < g.nn; n++) if(solid[n] == 0) cidx[(size_t)n] = c++;
    int64_t bad = 0;
    for(int64_t n = 0; n < g.nn; n++) {
        if(solid[n]) continue;
        const int64_t i = cidx[(size_t)n];
        const int x = (int)(n % g.nx);
        const int64_t row = n / g.nx;
        const int y = L::DIM == 3 ? (int)(row % g.ny) : (int)row;
        const int z = L::DIM == 3 ? (int)(row / g.ny) : 0;
        const CRec<NR> &R = tab[(size_t)i];
        if((int64_t)R.row != row) { bad++; continue; }
        // the kernel's decoding of the record
        uint32_t P[NR], start[NR], last[NR], off[L::Q];
        for(int m = 0; m < NR; m++) {
            start[m] = rowtab[(size_t)row * NW + m];
            last[m] = rowtab[(size_t)row * NW + NR + m];
        }
        P[0] = (uint32_t)i - start[0];
        for(int m = 1; m < NR; m++) P[m] = R.rank[m - 1];
        compact_offsets<L>((uint32_t)i, R.flags, P, start, last, off);
        if(off[0] != pop_at<L::Q>(0, (uint32_t)i)) bad++;
        for(int k = 1; k < L::Q; k++) {
            const int xx = (x - L::ex(k) + g.nx) % g.nx;
            const int yy = (y - L::ey(k) + g.ny) % g.ny;
            const int zz = L::DIM == 3 ? (z - L::ez(k) + g.nz) % g.nz : 0;
            const int64_t dn = ((int64_t)zz * g.ny + yy) * g.nx + xx;
            const uint32_t want = solid[(size_t)dn]
                                      ? pop_at<L::Q>(k, (uint32_t)i)
                                      : pop_at<L::Q>(L::opp(k), (uint32_t)cidx[(size_t)dn]);
            if(off[k] != want) bad++;
        }
    }
    return bad;
}

std::string geometry_verify_links(const uint8_t *solid, int dim, int nx, int ny, int nz, int64_t *bad)
{
    Geometry g;
    std::vector<SegInfo> h;
    std::string err = build_segments(g, solid, dim, nx, ny, nz, h);
    if(!err.empty()) return err;
    g.solid.assign(solid, solid + g.nn);
    *bad = dim == 3 ? verify_links<D3Q19>(g, solid, h) + verify_compact<D3Q19>(g, solid, h)
                    : verify_links<D2Q9>(g, solid, h) + verify_compact<D2Q9>(g, solid, h);
    return "";
}

std::string geometry_build_compact(Geometry &g)
{
    if(g.d_ctab) return "";
    std::vector<SegInfo> h((size_t)g.nunits);
    hipError_t e = hipMemcpy(h.data(), g.d_seg, h.size() * sizeof(SegInfo), hipMemcpyDeviceToHost);
    if(e != hipSuccess) return hip_err("hipMemcpy(segments)", e);
    std::vector<uint32_t> rowbase;
    size_t tab_bytes = 0;
    const void *tab_data = nullptr;
    std::vector<CRec<RowSlots<D3Q19>::N>> t3;
    std::vector<CRec<RowSlots<D2Q9>::N>> t2;
    bool ok;
    if(g.dim == 3) { ok = build_compact<D3Q19>(g, h, t3, rowbase); tab_bytes = t3.size() * sizeof(t3[0]); tab_data = t3.data(); }
    else           { ok = build_compact<D2Q9>(g, h, t2, rowbase); tab_bytes = t2.size() * sizeof(t2[0]); tab_data = t2.data(); }
    if(!ok) return "rows too long for the compact neighbour table";
    e = hipMalloc(&g.d_ctab, tab_bytes);
    if(e != hipSuccess) { g.d_ctab = nullptr; (void)hipGetLastError(); return "out of device memory"; }
    e = hipMalloc((void **)&g.d_rowbase, rowbase.size() * sizeof(uint32_t));
    if(e != hipSuccess) {
        (void)hipFree(g.d_ctab); g.d_ctab = nullptr; g.d_rowbase = nullptr;
        (void)hipGetLastError();
        return "out of device memory";
    }
    e = hipMemcpy(g.d_ctab, tab_data, tab_bytes, hipMemcpyHostToDevice);
    if(e == hipSuccess) e = hipMemcpy(g.d_rowbase, rowbase.data(), rowbase.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if(e != hipSuccess) return hip_err("hipMemcpy(compact table)", e);
    g.ctab_bytes = tab_bytes + rowbase.size() * sizeof(uint32_t);
    return "";
}

static int lattice_q(int lattice) { return lattice == 3 ? 19 : 9; }
static size_t elem_bytes(int prec) { return prec == PREC_F64 ? 8 : 4; }
static int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }
// population arrays are padded to whole 1024-cell blocks (the widest local-kernel workgroup)
static int64_t pop_stride(int64_t n_fluid) { return round_up(n_fluid, 1024); }

size_t job_bytes(const Geometry &g, int lattice, int prec, bool fields)
{
    const int64_t stride = pop_stride(g.n_fluid);
    size_t b = (size_t)lattice_q(lattice) * stride * elem_bytes(prec);
    if(fields) b += 2 * (size_t)g.dim * g.n_fluid * sizeof(float);
    return b;
}

size_t job_bytes_for(int64_t n_fluid, int dim, int lattice, int prec, bool fields)
{
    size_t b = (size_t)lattice_q(lattice) * pop_stride(n_fluid) * elem_bytes(prec);
    if(fields) b += 2 * (size_t)dim * n_fluid * sizeof(float);
    return b;
}

size_t Job::device_bytes() const { return geo ? job_bytes(*geo, lattice, prec, want_fields) : 0; }

void f32_force_terms(const double F[3], float g45[19], float h15[19])
{
    const float Ff[3] = {(float)F[0], (float)F[1], (float)F[2]};
    for(int k = 0; k < 19; k++) {
        const float w = (float)D3Q19::w(k);
        const float ef = D3Q19::ex(k) * Ff[0] + D3Q19::ey(k) * Ff[1] + D3Q19::ez(k) * Ff[2];
        g45[k] = 4.5f * w * ef;
        h15[k] = 1.5f * w * ef;
    }
}

static InstDesc make_desc(const Job &j)
{
    InstDesc D;
    std::memset(&D, 0, sizeof(D));
    const Geometry &g = *j.geo;
    D.seg = g.d_seg;
    D.links = g.d_links;
    D.rowspec = g.d_rowspec;
    D.ctab = g.d_ctab;
    D.rowbase = g.d_rowbase;
    D.f = j.d_f;
    D.vel[0] = j.d_vel[0];
    D.vel[1] = j.d_vel[1];
    D.stride = j.stride;
    D.n_fluid = g.n_fluid;
    D.nx = g.nx; D.ny = g.ny; D.nz = g.nz; D.nseg = g.nseg;
    for(int k = 0; k < 3; k++) { D.F[k] = j.F[k]; D.Ff[k] = (float)j.F[k]; }
    D.lattice = j.lattice;
    D.prec = j.prec;
    if(j.lattice == 3) f32_force_terms(j.F, D.g45, D.h15);
    return D;
}

int64_t pop_stride_of(int64_t n_fluid) { return pop_stride(n_fluid); }
InstDesc job_desc(const Job &j) { return make_desc(j); }

// ---------------------------------------------------------------------------

Batch::Batch(int lattice, int prec, hipStream_t stream, int chunk, bool use_graph)
    : lattice_(lattice), prec_(prec), stream_(stream), chunk_(chunk < 2 ? 2 : (chunk & ~1)),
      use_graph_(use_graph) {}

Batch::~Batch()
{
    (void)hipStreamSynchronize(stream_);
    release_graph();
    for(auto *list : {&jobs_, &pending_, &finished_})
        for(Job *j : *list) {
            if(!j) continue;
            free_job(j);
            if(j->admitted) { j->geo->refs--; j->admitted = false; }
        }
    (void)hipFree(d_inst_);
    (void)hipFree(d_cfg_);
    (void)hipFree(d_state_);
    (void)hipFree(d_state2_);
    (void)hipFree(d_lwork_);
    (void)hipFree(d_nwork_);
    (void)hipFree(d_cwork_);
    (void)hipFree(d_lpart_);
    (void)hipFree(d_npart_);
    (void)hipFree(d_lredw_);
    (void)hipFree(d_nredw_);
    (void)hipFree(d_lred_);
    (void)hipFree(d_nred_);
    (void)hipFree(d_count_);
    if(h_state_) (void)hipHostFree(h_state_);
}

void Batch::release_graph()
{
    for(int m = 0; m < 2; m++) {
        if(exec_[m]) { (void)hipGraphExecDestroy(exec_[m]); exec_[m] = nullptr; }
        if(graph_[m]) { (void)hipGraphDestroy(graph_[m]); graph_[m] = nullptr; }
    }
    for(PlanGraph &p : plan_graphs_) {
        if(p.x) (void)hipGraphExecDestroy(p.x);
        if(p.g) (void)hipGraphDestroy(p.g);
    }
    plan_graphs_.clear();
}

void Batch::free_job(Job *j)
{
    if(j->d_f) { (void)hipFree(j->d_f); j->d_f = nullptr; }
    for(int s = 0; s < 2; s++)
        if(j->d_vel[s]) { (void)hipFree(j->d_vel[s]); j->d_vel[s] = nullptr; }
}

std::string Batch::admit(Job *j)
{
    if(j->lattice != lattice_ || j->prec != prec_) return "job does not match the batch type";
    const Geometry &g = *j->geo;
    j->stride = pop_stride(g.n_fluid);
    const size_t fb = (size_t)lattice_q(lattice_) * j->stride * elem_bytes(prec_);
    hipError_t e = hipMalloc(&j->d_f, fb);
    if(e != hipSuccess) { j->d_f = nullptr; (void)hipGetLastError(); return "out of device memory"; }
    if(j->want_fields) {
        for(int s = 0; s < 2; s++) {
            e = hipMalloc(&j->d_vel[s], (size_t)g.dim * g.n_fluid * sizeof(float));
            if(e != hipSuccess) {
                j->d_vel[s] = nullptr;
                (void)hipGetLastError();
                free_job(j);
                return "out of device memory";
            }
        }
    }
    kern_rest(lattice_, prec_, make_desc(*j), stream_);
    e = hipGetLastError();
    if(e != hipSuccess) { free_job(j); return hip_err("rest-state kernel", e); }
    j->geo->refs++;
    init_state(j->cfg, j->state);
    j->admitted = true;
    j->done = j->state.phase >= PH_DONE;
    j->t_admit = now_seconds();
    if(j->done) { j->t_done = j->t_admit; finished_.push_back(j); }
    else { pending_.push_back(j); dirty_ = true; }
    return "";
}

// The host half of a table rebuild, for jobs whose neighbour-step form is
// decided (Job::nbr_mode): every kernel's work items, the flux-partial slots
// of each instance (InstDesc lp0..lp1 in the local partials, np0..np1 in the
// neighbour partials, where the row-segment items' slots come first and the
// compact items' follow), and the reduction items over those slots.  Sets each
// job's item count.
void build_tables(const std::vector<Job *> &jobs, int lattice, int prec, const Tuning &tu, BatchTables &t)
{
    const int n_inst = (int)jobs.size();
    std::vector<InstDesc> &desc = t.desc;
    std::vector<LocalWork> &lw = t.lw, &cw = t.cw;
    std::vector<NbrWork> &nw = t.nw;
    desc.assign((size_t)n_inst, InstDesc{});
    lw.clear(); cw.clear(); nw.clear();
    for(int i = 0; i < n_inst; i++) {
        Job *j = jobs[i];
        desc[i] = make_desc(*j);
        const Geometry &g = *j->geo;
        const int64_t lchunk = 256 * (int64_t)tu.local_cpt;
        desc[i].lp0 = (int32_t)lw.size();
        for(int64_t c = 0; c < g.n_fluid; c += lchunk) {
            LocalWork w;
            w.inst = i;
            w.ncell = (int32_t)std::min<int64_t>(lchunk, g.n_fluid - c);
            w.first = c;
            lw.push_back(w);
        }
        desc[i].lp1 = (int32_t)lw.size();
        // large lattices deal their neighbour-step items to the XCDs in y bands
        // (see below)
        const bool banded = g.dim == 3 &&
            (double)lattice_q(lattice) * (double)j->stride * (double)elem_bytes(prec) >= BAND_MIN_BYTES;
        if(j->nbr_mode == 1) {
            // compact neighbour step: up to 256 consecutive fluid cells per
            // item; its partial slots follow the row-segment ones (offset
            // fixed below).  Banded: the cells of rows y0..y1-1 of one plane
            // are consecutive, so band b is plane after plane of such runs
            // (256^3 FP32 0.385 against 0.406 ms, FP64 0.653 against 0.761;
            // 512^3 GRF FP32 0.700 against 0.712).
            desc[i].np0 = (int32_t)cw.size();
            auto runs = [&](int64_t a0, int64_t a1) {
                for(int64_t c = a0; c < a1; c += 256) {
                    LocalWork w;
                    w.inst = i;
                    w.ncell = (int32_t)std::min<int64_t>(256, a1 - c);
                    w.first = c;
                    cw.push_back(w);
                }
            };
            if(banded) {
                const int64_t band = (g.ny + 7) / 8;
                for(int64_t y0 = 0; y0 < g.ny; y0 += band) {
                    const int64_t y1 = std::min<int64_t>(g.ny, y0 + band);
                    for(int64_t z = 0; z < g.nz; z++)
                        runs(g.row_first[(size_t)(z * g.ny + y0)], g.row_first[(size_t)(z * g.ny + y1)]);
                }
            } else {
                runs(0, g.n_fluid);
            }
            desc[i].np1 = (int32_t)cw.size();
            continue;
        }
        desc[i].np0 = (int32_t)nw.size();
        {
            // Items of S x-adjacent segments of nbr_waves / S rows stacked
            // along z (2D: y).  A row's segments in one workgroup share the
            // lines at their boundaries in its L1: at 256^3 FP32 a whole row
            // per workgroup (S = 4) takes the neighbour step from 0.382 to
            // 0.345 ms against 4 z-stacked segments (S = 1).  Inside a slab y
            // then the segment vary fastest.  The kernels split their items
            // into 8 contiguous ranges, one per XCD.  Large lattices deal the
            // rows in 8 bands of ny/8 rows, each band through all z, slab after
            // slab: an XCD then owns one band and walks it in z, so the rows
            // either side of a slab boundary are in flight together and share
            // its L2 (256^3 FP32 step 0.318 against 0.332 ms; bands of 8-64
            // rows inside per-XCD z ranges gave 0.327-0.331).  Below ~400 MB of
            // populations the plain z order stays: an XCD's neighbour-step
            // range then covers the same cells as its local-step range, and
            // the L2 lines the one kernel leaves are the other's (192^3 FP32:
            // bands 0.139, z order 0.135 ms; 96^3: 0.030 against 0.027).
            // S = 2^sh x-adjacent segments per row (segs_per_item) and
            // nbr_waves / S rows per item
            int sh = 0;
            while((2 << sh) <= segs_per_item(tu, g.nseg)) sh++;
            const int64_t nwv = std::max(1, tu.nbr_waves >> sh);
            const int64_t planes = g.dim == 3 ? g.nz : g.ny;       // stacking axis
            const int64_t ys = g.dim == 3 ? g.ny : 1;
            const int64_t band = banded ? (ys + 7) / 8 : ys;
            for(int64_t y0 = 0; y0 < ys; y0 += band)
            for(int64_t p0 = 0; p0 < planes; p0 += nwv) {
                const int nu = (int)std::min<int64_t>(nwv, planes - p0);
                for(int64_t y = y0; y < std::min<int64_t>(ys, y0 + band); y++)
                    for(int sg = 0; sg < g.nseg; sg += 1 << sh) {
                        NbrWork w;
                        w.inst = i;
                        w.set_shape(sg, nu, sh);
                        w.y = (int32_t)y;
                        w.p0 = (int32_t)p0;
                        nw.push_back(w);
                    }
            }
        }
        desc[i].np1 = (int32_t)nw.size();
    }
    for(int i = 0; i < n_inst; i++) {
        if(jobs[i]->nbr_mode == 1) {
            desc[i].np0 += (int32_t)nw.size();
            desc[i].np1 += (int32_t)nw.size();
        }
        jobs[i]->items = (int64_t)(desc[i].lp1 - desc[i].lp0) + (desc[i].np1 - desc[i].np0);
    }
    // second-level reduction items over each instance's partial ranges
    std::vector<RedWork> &lr = t.lr, &nr = t.nr;
    lr.clear(); nr.clear();
    for(int i = 0; i < n_inst; i++) {
        auto split = [&](std::vector<RedWork> &out, int32_t p0, int32_t p1, int32_t &r0, int32_t &r1) {
            r0 = (int32_t)out.size();
            for(int32_t p = p0; p < p1; p += REDUCE_SPAN) {
                RedWork r;
                r.inst = i;
                r.p0 = p;
                r.p1 = std::min<int32_t>(p1, p + REDUCE_SPAN);
                r.pad = 0;
                out.push_back(r);
            }
            r1 = (int32_t)out.size();
        };
        split(lr, desc[i].lp0, desc[i].lp1, desc[i].rl0, desc[i].rl1);
        split(nr, desc[i].np0, desc[i].np1, desc[i].rn0, desc[i].rn1);
    }
}

// The flux-partial layout of a table set, as the kernels use it: local item
// it writes local slot it; row-segment item it writes neighbour slot it;
// compact item it writes neighbour slot n_rowseg + it (both partial arrays are
// allocated with exactly that many slots).  Checks that every slot an item
// writes lies in its array and in its own instance's range, that the ranges
// of the instances partition each array, and that each instance's reduction
// items tile its range in order, each at most REDUCE_SPAN slots, and are the
// ones its InstDesc names.  Returns the number of violations (the first one
// described in *first).
int segs_per_item(const Tuning &tu, int64_t nseg)
{
    // an item with segments past the row's end idles wavefronts: rows of 3
    // segments took 0.140 ms with S = 2 against 0.111 with S = 1 at 160^3
    const int cap = std::min(tu.nbr_segs, tu.nbr_waves);
    int s = 1;
    while(2 * s <= cap && nseg % (2 * s) == 0) s *= 2;
    return s;
}

int64_t check_tables(const BatchTables &t, std::string *first, int nbr_waves)
{
    int64_t bad = 0;
    auto fail = [&](const std::string &m) { if(!bad && first) *first = m; bad++; };
    // every (row, segment) unit of a row-segment instance in exactly one
    // wavefront of one item, whose workgroup has that wavefront
    {
        std::vector<std::vector<uint8_t>> seen(t.desc.size());
        for(size_t it = 0; it < t.nw.size(); it++) {
            const NbrWork &w = t.nw[it];
            if(w.inst < 0 || w.inst >= (int32_t)t.desc.size()) continue;          // reported below
            const InstDesc &D = t.desc[(size_t)w.inst];
            const int64_t nrows = D.lattice == 3 ? (int64_t)D.ny * D.nz : D.ny;
            std::vector<uint8_t> &sv = seen[(size_t)w.inst];
            if(sv.empty()) sv.assign((size_t)(nrows * D.nseg), 0);
            const int S = 1 << w.sshift();
            if((w.nunit() << w.sshift()) > nbr_waves || w.nunit() < 1)
                fail("row-segment item " + std::to_string(it) + " needs more wavefronts than a workgroup has");
            for(int dz = 0; dz < w.nunit(); dz++)
                for(int dsg = 0; dsg < S; dsg++) {
                    const int64_t sg = w.seg() + dsg, p = w.p0 + dz;
                    if(sg >= D.nseg) continue;                                   // an idle wavefront
                    const int64_t row = D.lattice == 3 ? p * D.ny + w.y : p;
                    if(row < 0 || row >= nrows || sg < 0 || (D.lattice == 3 && (w.y < 0 || w.y >= D.ny))) {
                        fail("row-segment item " + std::to_string(it) + " outside its lattice");
                        continue;
                    }
                    if(sv[(size_t)(row * D.nseg + sg)]++)
                        fail("unit of row " + std::to_string(row) + " segment " + std::to_string(sg) + " stepped twice");
                }
        }
        for(size_t i = 0; i < seen.size(); i++)
            for(size_t u = 0; u < seen[i].size(); u++)
                if(!seen[i][u]) { fail("instance " + std::to_string(i) + " unit " + std::to_string(u) + " in no item"); break; }
    }
    const int64_t n_inst = (int64_t)t.desc.size();
    const int64_t nl = (int64_t)t.lw.size(), nn = (int64_t)t.nw.size() + (int64_t)t.cw.size();
    auto in_inst = [&](int64_t slot, int32_t inst, bool local) {
        if(inst < 0 || inst >= n_inst) return false;
        const InstDesc &D = t.desc[(size_t)inst];
        return local ? (slot >= D.lp0 && slot < D.lp1) : (slot >= D.np0 && slot < D.np1);
    };
    for(int64_t it = 0; it < nl; it++)
        if(!in_inst(it, t.lw[(size_t)it].inst, true)) fail("local item " + std::to_string(it) + " outside its instance");
    for(int64_t it = 0; it < (int64_t)t.nw.size(); it++)
        if(!in_inst(it, t.nw[(size_t)it].inst, false)) fail("row-segment item " + std::to_string(it) + " outside its instance");
    for(int64_t it = 0; it < (int64_t)t.cw.size(); it++) {
        const int64_t slot = (int64_t)t.nw.size() + it;
        if(slot >= nn || !in_inst(slot, t.cw[(size_t)it].inst, false))
            fail("compact item " + std::to_string(it) + " outside its instance");
    }
    for(int local = 0; local < 2; local++) {
        const int64_t size = local ? nl : nn;
        const std::vector<RedWork> &rw = local ? t.lr : t.nr;
        std::vector<int> owner((size_t)size, -1);
        for(int64_t i = 0; i < n_inst; i++) {
            const InstDesc &D = t.desc[(size_t)i];
            const int32_t p0 = local ? D.lp0 : D.np0, p1 = local ? D.lp1 : D.np1;
            const int32_t r0 = local ? D.rl0 : D.rn0, r1 = local ? D.rl1 : D.rn1;
            if(p0 < 0 || p1 > size || p0 > p1) { fail("instance " + std::to_string(i) + " range outside the array"); continue; }
            for(int32_t p = p0; p < p1; p++) {
                if(owner[(size_t)p] >= 0) fail("slot " + std::to_string(p) + " in two instances");
                owner[(size_t)p] = (int)i;
            }
            int32_t next = p0;
            if(r0 < 0 || r1 > (int32_t)rw.size() || r0 > r1) { fail("reduction items outside the list"); continue; }
            for(int32_t r = r0; r < r1; r++) {
                const RedWork &w = rw[(size_t)r];
                if(w.inst != i || w.p0 != next || w.p1 <= w.p0 || w.p1 - w.p0 > REDUCE_SPAN)
                    fail("reduction item " + std::to_string(r) + " does not tile instance " + std::to_string(i));
                next = w.p1;
            }
            if(next != p1) fail("reduction items of instance " + std::to_string(i) + " stop short");
        }
        for(int64_t p = 0; p < size; p++)
            if(owner[(size_t)p] < 0) fail("slot " + std::to_string(p) + " in no instance");
    }
    return bad;
}

// Host only: the tables a batch would build for these structures' forcing
// runs (3 per 3D structure, 2 per 2D one; neighbour-step forms chosen as the
// batch chooses them, without building compact tables), checked.
std::string verify_batch_tables(const std::vector<const uint8_t *> &solids,
                                const std::vector<int> &dims, int lattice, int prec,
                                const Tuning &tu, int64_t *bad)
{
    std::vector<std::unique_ptr<Geometry>> geos;
    std::vector<std::unique_ptr<Job>> jobs;
    std::vector<Job *> ptrs;
    const int dim = lattice == 3 ? 3 : 2;
    for(size_t s = 0; s < solids.size(); s++) {
        std::unique_ptr<Geometry> g(new Geometry);
        std::vector<SegInfo> h;
        const int nx = dims[3 * s], ny = dims[3 * s + 1], nz = dim == 3 ? dims[3 * s + 2] : 1;
        std::string err = build_segments(*g, solids[s], dim, nx, ny, nz, h);
        if(!err.empty()) return err;
        g->solid.assign(solids[s], solids[s] + g->nn);
        g->row_first.resize((size_t)g->nrows + 1);
        for(int64_t r = 0; r < g->nrows; r++) g->row_first[(size_t)r] = h[(size_t)(r * g->nseg)].start;
        g->row_first[(size_t)g->nrows] = (uint32_t)g->n_fluid;
        for(int d = 0; d < dim; d++) {
            std::unique_ptr<Job> j(new Job);
            j->geo = g.get();
            j->lattice = lattice;
            j->prec = prec;
            j->stride = pop_stride(g->n_fluid);
            int m = tu.nbr_mode;
            if(m == 2)
                m = prefer_compact((double)g->n_fluid / (double)g->nn, segs_per_item(tu, g->nseg),
                                   lattice, prec) ? 1 : 0;
            if(m == 1 && g->nx >= 0xFFFF) m = 0;
            j->nbr_mode = m;
            ptrs.push_back(j.get());
            jobs.push_back(std::move(j));
        }
        geos.push_back(std::move(g));
    }
    BatchTables t;
    build_tables(ptrs, lattice, prec, tu, t);
    std::string first;
    *bad = check_tables(t, &first, tu.nbr_waves);
    return first;
}

static std::atomic<int> g_fail_rebuilds{0};

void test_fail_rebuilds(int n) { g_fail_rebuilds.store(n < 0 ? 0 : n); }

template<class T>
static hipError_t ensure(T *&p, size_t &cap, size_t n)
{
    if(n <= cap && p) return hipSuccess;
    if(p) (void)hipFree(p);
    p = nullptr;
    size_t c = std::max<size_t>(n, cap * 3 / 2 + 1);
    hipError_t e = hipMalloc(&p, c * sizeof(T));
    if(e == hipSuccess) cap = c; else { p = nullptr; cap = 0; }
    return e;
}

std::string Batch::rebuild()
{
    for(int left = g_fail_rebuilds.load(); left > 0; left = g_fail_rebuilds.load())
        if(g_fail_rebuilds.compare_exchange_weak(left, left - 1)) return "injected device failure";
    hipError_t e = hipStreamSynchronize(stream_);
    if(e != hipSuccess) return hip_err("stream", e);
    release_graph();
    // drop finished jobs from the table, take in the pending ones
    std::vector<Job *> live;
    for(Job *j : jobs_) if(j && !j->done) live.push_back(j);
    for(Job *j : pending_) live.push_back(j);
    pending_.clear();
    jobs_ = live;
    n_inst_ = (int)jobs_.size();
    work_total_ = work_dead_ = 0;
    if(n_inst_ == 0) { dirty_ = false; n_lwork_ = n_nwork_ = 0; return ""; }

    // neighbour-step form of every instance: the tuning's, or by porosity
    // (auto); the compact form needs the structure's compact table, and an
    // instance whose structure cannot have it (rows of 0xFFFF+ cells, or no
    // memory) uses the row-segment form
    int n_mode[2] = {0, 0};
    for(Job *j : jobs_) {
        const Geometry &g = *j->geo;
        int m = tune_.nbr_mode;
        if(m == 2)
            m = prefer_compact((double)g.n_fluid / (double)g.nn, segs_per_item(tune_, g.nseg),
                               lattice_, prec_) ? 1 : 0;
        if(m == 1 && !geometry_build_compact(*j->geo).empty()) m = 0;
        j->nbr_mode = m;
        n_mode[m]++;
    }
    mode_ = n_mode[1] == 0 ? 0 : (n_mode[0] == 0 ? 1 : 2);
    // the narrowest population addressing that covers every instance:
    // 32-bit byte offsets below 4 GiB, 32-bit element indices below 2^32
    // elements, 64-bit element indices beyond.  FLOWPERM_WIDE_OFFSETS=1 or 2
    // forces at least the element-index or the 64-bit kernels, so tests can
    // compare them with the narrow ones on small lattices.
    const char *wide = getenv("FLOWPERM_WIDE_OFFSETS");
    ad_ = AD_B32;
    if(wide && wide[0] == '1') ad_ = AD_E32;
    if(wide && wide[0] == '2') ad_ = AD_E64;
    for(const Job *j : jobs_) {
        const uint64_t elems = (uint64_t)lattice_q(lattice_) * (uint64_t)j->stride;
        if(elems * elem_bytes(prec_) > 0xFFFFFFFFull) ad_ = std::max(ad_, (int)AD_E32);
        if(elems > 0xFFFFFFFFull) ad_ = AD_E64;
    }
    BatchTables tab;
    build_tables(jobs_, lattice_, prec_, tune_, tab);
    std::vector<InstDesc> &desc = tab.desc;
    std::vector<LocalWork> &lw = tab.lw, &cw = tab.cw;
    std::vector<NbrWork> &nw = tab.nw;
    std::vector<RedWork> &lr = tab.lr, &nr = tab.nr;
    std::vector<InstCfg> cfg(n_inst_);
    std::vector<InstState> state(n_inst_);
    for(int i = 0; i < n_inst_; i++) { cfg[i] = jobs_[i]->cfg; state[i] = jobs_[i]->state; }
    n_lred_ = (int)lr.size();
    n_nred_ = (int)nr.size();
    // the deferred finish has every workgroup of an instance sum the
    // instance's partials of the previous step: up to REDUCE_SPAN of them on
    // any step, up to DEFER_MAX_PARTIALS when the flux is summed on few steps
    // (choose_defer); FLOWPERM_DEFER_FINISH=0 turns it off
    const char *sv = getenv("FLOWPERM_SPARSE_FINISH");      // 0: a finish after every step
    sparse_ok_ = !(sv && sv[0] == '0');
    const char *dv = getenv("FLOWPERM_DEFER_FINISH");
    defer_ok_ = !(dv && dv[0] == '0') && ad_ == AD_B32;
    defer_small_ = defer_ok_;
    for(int i = 0; i < n_inst_; i++) {
        const int32_t nl = desc[i].lp1 - desc[i].lp0, nn = desc[i].np1 - desc[i].np0;
        if(nl > DEFER_MAX_PARTIALS || nn > DEFER_MAX_PARTIALS) defer_ok_ = false;
        if(nl > REDUCE_SPAN || nn > REDUCE_SPAN) defer_small_ = false;
    }
    defer_small_ = defer_small_ && defer_ok_;
    if(lw.size() > INT32_MAX || nw.size() > INT32_MAX || cw.size() > INT32_MAX) return "batch too large";
    n_lwork_ = (int)lw.size();
    n_nwork_ = (int)nw.size();
    n_cwork_ = (int)cw.size();
    work_total_ = (int64_t)lw.size() + (int64_t)n_nwork_ + (int64_t)n_cwork_;
    const size_t ni = (size_t)n_inst_;
    if(ni > cap_inst_ || !d_inst_) {
        (void)hipFree(d_inst_); (void)hipFree(d_cfg_); (void)hipFree(d_state_); (void)hipFree(d_state2_);
        d_inst_ = nullptr; d_cfg_ = nullptr; d_state_ = nullptr; d_state2_ = nullptr;
        if(h_state_) { (void)hipHostFree(h_state_); h_state_ = nullptr; }
        cap_inst_ = 0;
        const size_t c = std::max(ni, (size_t)16);
        if((e = hipMalloc(&d_inst_, c * sizeof(InstDesc))) != hipSuccess ||
           (e = hipMalloc(&d_cfg_, c * sizeof(InstCfg))) != hipSuccess ||
           (e = hipMalloc(&d_state_, c * sizeof(InstState))) != hipSuccess ||
           (e = hipMalloc(&d_state2_, c * sizeof(InstState))) != hipSuccess ||
           (e = hipHostMalloc((void **)&h_state_, c * sizeof(InstState), hipHostMallocDefault)) != hipSuccess)
            return hip_err("allocating batch tables", e);
        cap_inst_ = c;
    }
    if((e = ensure(d_lwork_, cap_lwork_, std::max<size_t>(1, lw.size()))) != hipSuccess) return hip_err("hipMalloc(work)", e);
    if((e = ensure(d_nwork_, cap_nwork_, std::max<size_t>(1, nw.size()))) != hipSuccess) return hip_err("hipMalloc(work)", e);
    if((e = ensure(d_cwork_, cap_cwork_, std::max<size_t>(1, cw.size()))) != hipSuccess) return hip_err("hipMalloc(work)", e);
    if((e = ensure(d_lpart_, cap_lpart_, std::max<size_t>(1, lw.size()))) != hipSuccess) return hip_err("hipMalloc(partials)", e);
    if((e = ensure(d_npart_, cap_npart_, std::max<size_t>(1, (size_t)n_nwork_ + n_cwork_))) != hipSuccess) return hip_err("hipMalloc(partials)", e);
    if((e = ensure(d_lredw_, cap_lredw_, std::max<size_t>(1, lr.size()))) != hipSuccess ||
       (e = ensure(d_nredw_, cap_nredw_, std::max<size_t>(1, nr.size()))) != hipSuccess ||
       (e = ensure(d_lred_, cap_lred_, std::max<size_t>(1, lr.size()))) != hipSuccess ||
       (e = ensure(d_nred_, cap_nred_, std::max<size_t>(1, nr.size()))) != hipSuccess ||
       (e = ensure(d_count_, cap_count_, (size_t)n_inst_)) != hipSuccess)
        return hip_err("hipMalloc(reduction)", e);
    e = hipMemcpy(d_inst_, desc.data(), desc.size() * sizeof(InstDesc), hipMemcpyHostToDevice);
    if(e == hipSuccess) e = hipMemcpy(d_cfg_, cfg.data(), cfg.size() * sizeof(InstCfg), hipMemcpyHostToDevice);
    if(e == hipSuccess) e = hipMemcpy(d_state_, state.data(), state.size() * sizeof(InstState), hipMemcpyHostToDevice);
    // the deferred finish's second state array is written by every step before
    // it is read; it starts as a copy so that no kernel can ever see garbage
    if(e == hipSuccess) e = hipMemcpy(d_state2_, state.data(), state.size() * sizeof(InstState), hipMemcpyHostToDevice);
    if(e == hipSuccess && !lw.empty()) e = hipMemcpy(d_lwork_, lw.data(), lw.size() * sizeof(LocalWork), hipMemcpyHostToDevice);
    if(e == hipSuccess && !nw.empty()) e = hipMemcpy(d_nwork_, nw.data(), nw.size() * sizeof(NbrWork), hipMemcpyHostToDevice);
    if(e == hipSuccess && !cw.empty()) e = hipMemcpy(d_cwork_, cw.data(), cw.size() * sizeof(LocalWork), hipMemcpyHostToDevice);
    if(e == hipSuccess && !lr.empty()) e = hipMemcpy(d_lredw_, lr.data(), lr.size() * sizeof(RedWork), hipMemcpyHostToDevice);
    if(e == hipSuccess && !nr.empty()) e = hipMemcpy(d_nredw_, nr.data(), nr.size() * sizeof(RedWork), hipMemcpyHostToDevice);
    if(e == hipSuccess) {
        // the finish kernel's counters start at zero (and it leaves them so)
        const std::vector<unsigned int> zero((size_t)n_inst_, 0u);
        e = hipMemcpy(d_count_, zero.data(), zero.size() * sizeof(unsigned int), hipMemcpyHostToDevice);
    }
    if(e != hipSuccess) return hip_err("hipMemcpy(tables)", e);
    dirty_ = false;
    return "";
}

bool Batch::wants_rebuild() const
{
    // jobs that retired stay in the tables (the kernels skip them after one
    // load of their phase) until a quarter of the launched work items are
    // theirs, so that staggered retirements do not each cost a new graph
    return dirty_ || !pending_.empty() || (work_total_ > 0 && 4 * work_dead_ > work_total_);
}

// Deferred finish for the next chunk?  Always when every instance has at most
// REDUCE_SPAN partials per step kernel; with up to DEFER_MAX_PARTIALS only
// while no live instance sums its flux on most steps (a 3D averaging phase,
// or checks every few steps), where every workgroup summing thousands of
// partials after every step would cost more than the finish launch it saves.
bool Batch::choose_defer() const
{
    if(!defer_ok_) return false;
    if(defer_small_) return true;
    for(const Job *j : jobs_) {
        if(!j || j->done) continue;
        const InstCfg &c = j->cfg;
        const bool dense = c.kind == 3 ? (j->state.phase == PH_AVG || c.lag < 16) : c.avg_sample < 16;
        if(dense) return false;
    }
    return true;
}

std::string Batch::run_chunk()
{
    if(wants_rebuild()) {
        std::string err = rebuild();
        if(!err.empty()) return err;
    }
    if(n_inst_ == 0) return "";
    const double t0 = now_seconds();
    hipError_t e;
    // the chunk's finish plan, or (null) the deferred finish where it
    // applies, else a finish after every step
    std::vector<int> adv;
    bool can_defer = false;
    const std::vector<int> *plan = choose_plan(0, chunk_, 1, adv, can_defer) ? &adv : nullptr;
    auto capture = [&](const std::vector<int> *pl, hipGraph_t &g, hipGraphExec_t &x) -> std::string {
        hipError_t ce = hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal);
        if(ce != hipSuccess) return hip_err("hipStreamBeginCapture", ce);
        enqueue_steps(chunk_, pl);
        ce = hipStreamEndCapture(stream_, &g);
        if(ce != hipSuccess) { g = nullptr; return hip_err("hipStreamEndCapture", ce); }
        ce = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        if(ce != hipSuccess) { x = nullptr; return hip_err("hipGraphInstantiate", ce); }
        return "";
    };
    if(use_graph_) {
        hipGraphExec_t x = nullptr;
        defer_ = false;
        if(plan) {
            for(const PlanGraph &pg : plan_graphs_)
                if(pg.adv == adv) { x = pg.x; break; }
            // A job set's plans repeat with the lags, so a few graphs cover
            // them; lags that do not divide the chunk can give more plans than
            // that, and a full cache then runs the chunk without a plan rather
            // than capturing a graph per chunk.
            if(!x && plan_graphs_.size() < MAX_PLAN_GRAPHS) {
                PlanGraph pg;
                pg.adv = adv;
                std::string err = capture(plan, pg.g, pg.x);
                if(!err.empty()) {
                    if(pg.g) (void)hipGraphDestroy(pg.g);
                    return err;
                }
                x = pg.x;
                plan_graphs_.push_back(std::move(pg));
            }
        }
        if(!x) {
            defer_ = can_defer;
            const int m = defer_ ? 1 : 0;
            if(!exec_[m]) {
                std::string err = capture(nullptr, graph_[m], exec_[m]);
                if(!err.empty()) return err;
            }
            x = exec_[m];
        }
        e = hipGraphLaunch(x, stream_);
        if(e != hipSuccess) return hip_err("hipGraphLaunch", e);
    } else {
        defer_ = !plan && can_defer;
        enqueue_steps(chunk_, plan);
        e = hipGetLastError();
        if(e != hipSuccess) return hip_err("kernel launch", e);
    }
    e = hipMemcpyAsync(h_state_, d_state_, n_inst_ * sizeof(InstState), hipMemcpyDeviceToHost, stream_);
    if(e != hipSuccess) return hip_err("hipMemcpyAsync(state)", e);
    e = hipStreamSynchronize(stream_);
    if(e != hipSuccess) return hip_err("device", e);
    stats_.steps += chunk_;
    stats_.chunks += 1;
    stats_.gpu_ms += (now_seconds() - t0) * 1e3;
    const double t1 = now_seconds();
    for(int i = 0; i < n_inst_; i++) {
        Job *j = jobs_[i];
        if(!j || j->done) continue;
        j->state = h_state_[i];
        if(j->state.phase >= PH_DONE) {
            j->done = true;
            j->t_done = t1;
            finished_.push_back(j);
            work_dead_ += work_items(*j);
        }
    }
    return "";
}

int64_t Batch::work_items(const Job &j) const
{
    return j.items;
}

void Batch::launch_finish(bool local, int adv)
{
    kern_finish(local ? n_lred_ : n_nred_, local ? d_lredw_ : d_nredw_, d_inst_, d_cfg_, d_state_,
                local ? d_lpart_ : d_npart_, local ? d_lred_ : d_nred_, d_count_, local, stream_, adv);
}

void Batch::plan_chunk(int64_t skip, int n, int replays, std::vector<int> &adv) const
{
    std::vector<InstCfg> cfg;
    std::vector<InstState> st;
    for(const Job *j : jobs_) {
        if(!j || j->done) continue;
        cfg.push_back(j->cfg);
        st.push_back(j->state);
    }
    plan_finishes(cfg.data(), st.data(), (int)cfg.size(), skip, n, replays, adv);
}

bool Batch::choose_plan(int64_t skip, int n, int replays, std::vector<int> &adv, bool &can_defer) const
{
    can_defer = choose_defer();
    if(!sparse_ok_) return false;
    plan_chunk(skip, n, replays, adv);
    const int64_t nfin = std::count_if(adv.begin(), adv.end(), [](int a) { return a > 0; });
    if(nfin >= n) return false;                      // a finish after every step: no plan
    // the deferred update costs every step kernel 0.7-2 us on the lattices it
    // takes, a finish launch 4-8 us: a plan wins while it keeps few finishes
    return !(can_defer && nfin * PLAN_OVER_DEFER > n);
}

void Batch::launch_step(bool local, const Defer *df)
{
    // the state the kernels read: the current one, or (deferred) the one
    // before the previous step's finish
    const InstState *st = d_state_;
    Defer d = {};
    if(df) {
        d = *df;
        st = d.sout == d_state2_ ? d_state_ : d_state2_;
        d.slot0 = 0;
    }
    // The workgroup caps pay where steps are mostly quiet (no flux sum): 256^3
    // FP32, 0.2875 against 0.2958 ms a step.  With a flux sum on every step each
    // work item ends in an LDS exchange and a barrier, which fewer resident
    // workgroups hide worse (0.342 against 0.324 ms), so chunks without a finish
    // plan (dense or deferred finish) launch uncapped.  D2Q9 keeps 9 populations
    // of a cell in flight, not 19, and runs slower capped (1024^2 FP32: 9.3
    // against 8.4 us a step): the caps are for D3Q19 only.
    Tuning tu = tune_;
    if(!capped_ || lattice_ != 3) tu.local_wgs = tu.nbr_wgs = 0;
    if(local) {
        kern_step(lattice_, prec_, true, tu, ad_, d_inst_, st, d_lwork_, n_lwork_, d_lpart_, stream_,
                  df ? &d : nullptr);
        return;
    }
    // the row-segment items, then the compact ones (their partial slots follow)
    if(n_nwork_ > 0) {
        tu.nbr_mode = 0;
        kern_step(lattice_, prec_, false, tu, ad_, d_inst_, st, d_nwork_, n_nwork_, d_npart_, stream_,
                  df ? &d : nullptr);
    }
    if(n_cwork_ > 0) {
        tu.nbr_mode = 1;
        d.slot0 = n_nwork_;
        kern_step(lattice_, prec_, false, tu, ad_, d_inst_, st, d_cwork_, n_cwork_,
                  d_npart_ + n_nwork_, stream_, df ? &d : nullptr);
    }
}

void Batch::enqueue_steps(int n, const std::vector<int> *adv)
{
    capped_ = adv != nullptr;
    if(!defer_) {
        for(int t = 0; t < n; t++) {
            const bool local = (t & 1) == 0;     // every live job starts at an even step count
            launch_step(local);
            const int a = adv ? (*adv)[t] : 1;
            if(a > 0) launch_finish(local, a);
        }
        return;
    }
    // step t reads one state array and writes the other: t = 0 reads the
    // current state (nothing pending) from d_state_ and hands it on to
    // d_state2_; every later step applies the previous step's finish first.
    // n is even, so the last step writes d_state_, and the finish launch of
    // the last (neighbour) step completes it there.
    for(int t = 0; t < n; t++) {
        const bool local = (t & 1) == 0;
        Defer df;
        df.cfg = d_cfg_;
        df.sout = (t & 1) == 0 ? d_state2_ : d_state_;
        df.prev = local ? d_npart_ : d_lpart_;
        df.pending = t > 0;
        df.slot0 = 0;
        launch_step(local, &df);
    }
    launch_finish(false);
}

std::string Batch::time_steps(int steps, KernelTimes &kt)
{
    if(wants_rebuild()) {
        std::string err = rebuild();
        if(!err.empty()) return err;
    }
    if(n_inst_ == 0) return "nothing to step";
    steps &= ~1;
    {
        // the caps where a graph of these steps would apply them
        std::vector<int> adv;
        bool cd = false;
        capped_ = choose_plan(0, steps, 1, adv, cd);
    }
    const int nev = 4 * steps + 2;
    std::vector<hipEvent_t> ev(nev);
    for(auto &x : ev)
        if(hipEventCreate(&x) != hipSuccess) return "hipEventCreate failed";
    int k = 0;
    (void)hipEventRecord(ev[k++], stream_);
    for(int t = 0; t < steps; t++) {
        const bool local = (t & 1) == 0;
        (void)hipEventRecord(ev[k++], stream_);
        launch_step(local);
        (void)hipEventRecord(ev[k++], stream_);
        launch_finish(local);
        (void)hipEventRecord(ev[k++], stream_);
        (void)hipEventRecord(ev[k++], stream_);
    }
    (void)hipEventRecord(ev[k++], stream_);
    hipError_t e = hipStreamSynchronize(stream_);
    if(e != hipSuccess) return hip_err("device", e);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev[0], ev[k - 1]);
    kt.total_ms = ms;
    for(int t = 0; t < steps; t++) {
        const int b = 1 + 4 * t;
        float a = 0, m = 0;
        (void)hipEventElapsedTime(&a, ev[b], ev[b + 1]);
        (void)hipEventElapsedTime(&m, ev[b + 1], ev[b + 2]);
        if((t & 1) == 0) { kt.local_ms += a; kt.n_local++; }
        else             { kt.nbr_ms += a; kt.n_nbr++; }
        kt.monitor_ms += m;
        kt.n_monitor++;
    }
    for(auto &x : ev) (void)hipEventDestroy(x);
    // pull the states back so the host mirror stays current
    e = hipMemcpy(h_state_, d_state_, n_inst_ * sizeof(InstState), hipMemcpyDeviceToHost);
    if(e != hipSuccess) return hip_err("hipMemcpy(state)", e);
    for(int i = 0; i < n_inst_; i++) if(jobs_[i]) jobs_[i]->state = h_state_[i];
    return "";
}

std::string Batch::graph_steps(int steps, float &ms)
{
    ms = 0.0f;
    if(wants_rebuild()) {
        std::string err = rebuild();
        if(!err.empty()) return err;
    }
    if(n_inst_ == 0) return "nothing to step";
    steps &= ~1;
    if(steps < 2) return "";
    const int c = std::min(steps, 200) & ~1;         // every replay starts at an even step
    const int reps = steps / c, rest = steps - reps * c;
    hipGraph_t g[2] = {nullptr, nullptr};
    hipGraphExec_t x[2] = {nullptr, nullptr};
    hipEvent_t a = nullptr, b = nullptr;
    // the modes of the two graphs: the chunk graph runs 1 + reps times (one
    // untimed), the remainder's once after them
    std::vector<int> adv[2];
    const std::vector<int> *plan[2] = {nullptr, nullptr};
    bool dfr[2] = {false, false};
    for(int k = 0; k < (rest ? 2 : 1); k++) {
        bool cd = false;
        if(choose_plan(k ? (int64_t)(1 + reps) * c : 0, k ? rest : c, k ? 1 : 1 + reps, adv[k], cd))
            plan[k] = &adv[k];
        else
            dfr[k] = cd;
    }
    auto capture = [&](int n, int k) -> std::string {
        defer_ = dfr[k];
        hipError_t e = hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal);
        if(e != hipSuccess) return hip_err("hipStreamBeginCapture", e);
        enqueue_steps(n, plan[k]);
        e = hipStreamEndCapture(stream_, &g[k]);
        if(e != hipSuccess) { g[k] = nullptr; return hip_err("hipStreamEndCapture", e); }
        e = hipGraphInstantiate(&x[k], g[k], nullptr, nullptr, 0);
        if(e != hipSuccess) { x[k] = nullptr; return hip_err("hipGraphInstantiate", e); }
        return "";
    };
    std::string err = capture(c, 0);
    if(err.empty() && rest) err = capture(rest, 1);
    defer_ = dfr[0];                                 // the mode of the timed replays
    if(err.empty() && (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess))
        err = "hipEventCreate failed";
    if(err.empty()) {
        hipError_t e = hipGraphLaunch(x[0], stream_);                // untimed
        if(e == hipSuccess) e = hipEventRecord(a, stream_);
        for(int r = 0; r < reps && e == hipSuccess; r++) e = hipGraphLaunch(x[0], stream_);
        if(e == hipSuccess && rest) e = hipGraphLaunch(x[1], stream_);
        if(e == hipSuccess) e = hipEventRecord(b, stream_);
        if(e == hipSuccess) e = hipStreamSynchronize(stream_);
        if(e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        if(e != hipSuccess) err = hip_err("graph replay", e);
    }
    if(a) (void)hipEventDestroy(a);
    if(b) (void)hipEventDestroy(b);
    for(int i = 0; i < 2; i++) {
        if(x[i]) (void)hipGraphExecDestroy(x[i]);
        if(g[i]) (void)hipGraphDestroy(g[i]);
    }
    if(!err.empty()) return err;
    hipError_t e = hipMemcpy(h_state_, d_state_, n_inst_ * sizeof(InstState), hipMemcpyDeviceToHost);
    if(e != hipSuccess) return hip_err("hipMemcpy(state)", e);
    for(int i = 0; i < n_inst_; i++) if(jobs_[i]) jobs_[i]->state = h_state_[i];
    return "";
}

// ---------------------------------------------------------------------------
// finish plans
// ---------------------------------------------------------------------------
// the first step t > S.s after which a first phase may end: max_steps, or the
// check (t % lag == 0) at which hold can first reach c.hold.  hold grows by at
// most one per check, and only at checks at or after min_steps (a check before
// them resets it); with c.hold <= 0 any check ends the phase.
static int64_t first_possible_end(const InstCfg &c, const InstState &S)
{
    int64_t t = std::max(S.s + 1, c.max_steps);
    const int64_t lim = (INT64_MAX / 4) / c.lag;
    const int64_t b = S.s / c.lag;                   // checks so far: steps lag, ..., b * lag
    int64_t k = -1;                                  // the check (index) that may end it
    if(c.hold <= 0 || S.hold >= c.hold) {
        k = b + 1;
    } else if(hold_can_grow(c)) {
        const int64_t m = c.min_steps > 0 ? (c.min_steps - 1) / c.lag + 1 : 0;   // ceil(min_steps / lag)
        k = std::max(b + 1, m) + ((int64_t)c.hold - S.hold) - 1;
    }
    if(k > 0 && k <= lim) t = std::min(t, k * c.lag);
    return t;
}

void plan_finishes(const InstCfg *cfg, const InstState *st, int n_inst, int64_t skip, int n,
                   int replays, std::vector<int> &adv)
{
    std::vector<uint8_t> need(n, 0);
    need[n - 1] = 1;                                 // a chunk leaves every state current
    for(int i = 0; i < n_inst; i++) {
        const InstCfg &c = cfg[i];
        const InstState &S = st[i];
        if(S.phase >= PH_DONE) continue;
        if(S.phase == PH_AVG) {
            if(c.kind == 3) { std::fill(need.begin(), need.end(), 1); break; }
            // 2D: a sample every avg_sample steps, the end at avg_steps
            for(int r = 0; r < replays; r++) {
                const int64_t a0 = S.tavg + skip + (int64_t)r * n;   // averaging steps done before this run
                if(a0 >= c.avg_steps) break;
                for(int j = 0; j < n; j++) {
                    const int64_t ta = a0 + j + 1;
                    if(!quiet_avg_step(c, ta)) { need[j] = 1; if(j > 0) need[j - 1] = 1; }
                    if(ta >= c.avg_steps) break;                     // done after this step
                }
            }
            continue;
        }
        const int64_t tend = first_possible_end(c, S);
        for(int r = 0; r < replays; r++) {
            const int64_t t0 = S.s + skip + (int64_t)r * n;      // steps done before this run
            for(int j = 0; j < n; j++) {
                const int64_t t = t0 + j + 1;
                if(t >= tend) { std::fill(need.begin() + std::max(j - 1, 0), need.end(), 1); break; }
                if(!quiet_run_step(c, t)) { need[j] = 1; if(j > 0) need[j - 1] = 1; }
            }
        }
    }
    adv.assign(n, 0);
    int run = 0;
    for(int j = 0; j < n; j++) {
        run++;
        if(need[j]) { adv[j] = run; run = 0; }
    }
}

int64_t verify_finish_plan(const std::vector<InstCfg> &cfg, const double *q, int64_t nsteps,
                           const int64_t *start, int chunk, int replays, int plant,
                           int64_t *finishes)
{
    const int n = (int)cfg.size();
    std::vector<InstState> dense(n), sparse(n);
    std::vector<uint8_t> joined(n, 0);
    int64_t problems = 0, nfin = 0;
    const double zero[3] = {0.0, 0.0, 0.0};
    // the flux instance i's step t (1-based) returns
    auto flux_at = [&](int i, int64_t t) { return q + ((int64_t)i * nsteps + (t - 1)) * 3; };
    const int64_t round = (int64_t)chunk * replays;
    for(int64_t b = 0; b + round <= nsteps; b += round) {
        std::vector<InstCfg> lc;
        std::vector<InstState> ls;
        for(int i = 0; i < n; i++) {
            if(!joined[i] && start[i] <= b) {       // jobs join between chunks
                init_state(cfg[i], dense[i]);
                sparse[i] = dense[i];
                joined[i] = 1;
            }
            if(joined[i] && sparse[i].phase < PH_DONE) { lc.push_back(cfg[i]); ls.push_back(sparse[i]); }
        }
        std::vector<int> adv;
        plan_finishes(lc.data(), ls.data(), (int)lc.size(), 0, chunk, replays, adv);
        if(plant) { adv.assign(chunk, 0); adv[chunk - 1] = chunk; }
        for(int r = 0; r < replays; r++)
            for(int j = 0; j < chunk; j++) {
                nfin += adv[j] > 0;
                for(int i = 0; i < n; i++) {
                    if(!joined[i]) continue;
                    InstState &D = dense[i], &P = sparse[i];
                    if(D.phase >= PH_DONE) continue;             // the kernels skip it
                    // what the step kernels read
                    if(P.phase != D.phase || P.flux_next != D.flux_next || P.vel_next != D.vel_next)
                        problems++;
                    advance_state(cfg[i], D, D.flux_next ? flux_at(i, D.s + 1) : zero, D.flux_next);
                    plan_next(cfg[i], D);
                    if(adv[j] > 0 && P.phase < PH_DONE) {
                        const int64_t t = P.s + (P.phase == PH_RUN ? adv[j] : 1);
                        if(t > D.s) { problems++; continue; }
                        advance_state_n(cfg[i], P, P.flux_next ? flux_at(i, t) : zero, P.flux_next, adv[j]);
                        plan_next(cfg[i], P);
                        if(std::memcmp(&P, &D, sizeof(D)) != 0) problems++;
                    }
                }
            }
        for(int i = 0; i < n; i++)
            if(joined[i] && std::memcmp(&sparse[i], &dense[i], sizeof(InstState)) != 0) problems++;
    }
    if(finishes) *finishes = nfin;
    return problems;
}

std::vector<Job *> Batch::take_done()
{
    std::vector<Job *> out;
    out.swap(finished_);
    return out;
}

void Batch::detach_all()
{
    (void)hipStreamSynchronize(stream_);
    for(auto *list : {&jobs_, &pending_, &finished_})
        for(Job *j : *list)
            if(j && j->admitted) { j->geo->refs--; j->admitted = false; }
    jobs_.clear();
    pending_.clear();
    finished_.clear();
    dirty_ = true;
}

void Batch::release(Job *j)
{
    // the tables may name the job until the next rebuild, but its phase is
    // final, so no kernel reads its lattice (and no chunk is in flight here:
    // run_chunk returns after its stream synchronisation)
    if(!j->done) dirty_ = true;                      // a live job leaves: rebuild
    free_job(j);
    if(j->admitted) { j->geo->refs--; j->admitted = false; }
    for(Job *&p : jobs_) if(p == j) p = nullptr;     // the caller may now delete it
    pending_.erase(std::remove(pending_.begin(), pending_.end(), j), pending_.end());
    finished_.erase(std::remove(finished_.begin(), finished_.end(), j), finished_.end());
}

} // namespace fp
