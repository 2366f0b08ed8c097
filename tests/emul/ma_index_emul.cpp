// CPU emulation of the tiled mass-assignment index logic of csrc/ma.hip: home
// tiles, halo boxes, halo_index, band enumeration, axis_images, and the
// uniform-tile finalize.  Uses the same device helpers (compiled for the host)
// and checks, for every scheme, 2D and 3D, many grid sizes and edge positions:
//   - every LDS box index, partial slot and grid node is in bounds;
//   - every partial slot is written once and read only after it is written;
//   - every grid node is written exactly once;
//   - the result equals a direct float64 deposit;
//   - the uniform-tile finalize gives the same bits as the general one.
// Built and run by tests/test_ma_index_emul.py (CPU only, no GPU calls).
#include "../../pylians3_amd/csrc/ma.hip"
#include <cassert>
#include <random>

using namespace pyl;
static long long uni_checked = 0;

template <int MAS, int AXES>
static int run_case(int dims, int n, int pattern, bool weighted, unsigned seed) {
    constexpr int F = MAS;
    constexpr int T = (AXES == 3) ? kT3 : kT2;
    constexpr int BY = T + F - 1;
    constexpr int BZ = (AXES == 3) ? T + F - 1 : 1;
    constexpr int BX = T + F - 1;
    constexpr int NB = BX * BY * BZ;
    const float box = 100.0f;
    const float inv = (float)dims / box;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> U(0.0f, box);
    std::normal_distribution<float> Nd(0.0f, 1.0f);
    std::vector<float> pos((size_t)n * AXES), W(n, 1.0f);
    for (int i = 0; i < n; ++i) {
        for (int a = 0; a < AXES; ++a) {
            float x;
            if (pattern == 0) x = U(rng);
            else if (pattern == 1) {   // lattice-ish + displacement
                int side = (int)std::ceil(std::pow((double)n, 1.0 / AXES));
                int c = (a == 0) ? i / (side * (AXES == 3 ? side : 1)) : (a == 1 ? (i / (AXES == 3 ? side : 1)) % side : i % side);
                x = (c + 0.5f) * box / side + 3.0f * box / dims * Nd(rng);
                x = std::fmod(x, box); if (x < 0) x += box; if (x >= box) x = 0.0f;
            } else {                     // edges: nodes, half nodes, box ends
                int k = rng() % (dims + 1);
                float fr[6] = {0.0f, 0.5f, 0.25f, 0.999999f, 0.5000001f, 0.4999999f};
                x = (k + fr[rng() % 6]) * box / dims;
                if (rng() % 7 == 0) x = std::nextafter(box, 0.0f);
                if (x >= box) x = 0.0f;
            }
            pos[(size_t)i * AXES + a] = x;
        }
        if (weighted) W[i] = std::exp(Nd(rng));
    }
    TileGeom g = make_geom(dims, AXES, MAS);
    std::vector<int> orig(g.nt + 1);
    for (int i = 0; i <= g.nt; ++i) orig[i] = tile_origin(i, g.nt, dims);
    g.orig = orig.data();
    // direct float64 deposit
    const long long ncell = g.ncell;
    std::vector<double> ref(ncell, 0.0);
    for (int i = 0; i < n; ++i) {
        float C[3][4];
        int b[3] = {0, 0, 0};
        for (int a = 0; a < AXES; ++a) {
            float dist = pos[(size_t)i * AXES + a] * inv;
            int raw = first_node<MAS>(dist);
            foot_weights<MAS>(dist, raw, C[a]);
            b[a] = wrap_node(raw, dims);
        }
        for (int l = 0; l < F; ++l)
            for (int m = 0; m < F; ++m)
                for (int k = 0; k < (AXES == 3 ? F : 1); ++k) {
                    long long node = AXES == 3
                        ? ((long long)node_add(b[0], l, dims) * dims + node_add(b[1], m, dims)) * dims + node_add(b[2], k, dims)
                        : (long long)node_add(b[0], l, dims) * dims + node_add(b[1], m, dims);
                    double c = (double)C[0][l] * C[1][m] * (AXES == 3 ? C[2][k] : 1.0f) * W[i];
                    ref[node] += c;
                }
    }
    // keys and records (the sort order inside a tile does not matter)
    std::vector<std::vector<int>> bucket(g.ntiles);
    for (int i = 0; i < n; ++i) {
        int t[3] = {0, 0, 0};
        for (int a = 0; a < AXES; ++a) t[a] = g.shift >= 0 ? home_tile<MAS, true>(pos[(size_t)i * AXES + a] * inv, g) : home_tile<MAS, false>(pos[(size_t)i * AXES + a] * inv, g);
        long long key = ((long long)t[0] * g.nt + t[1]) * g.ntz + t[2];
        assert(key >= 0 && key < g.ntiles);
        bucket[key].push_back(i);
    }
    std::vector<double> grid(ncell, 0.0);
    std::vector<int> writes(ncell, 0);
    std::vector<double> part((size_t)g.ntiles * g.pslot, -1e300);
    for (long long key = 0; key < g.ntiles; ++key) {
        TileBox tb = tile_box<AXES>((int)key, g);
        assert(tb.ex >= 1 && tb.ex <= T && tb.ey >= 1 && tb.ey <= T && tb.ez >= 1);
        assert((halo_count<F, AXES>(tb.ex, tb.ey, tb.ez)) <= g.pslot);
        std::vector<double> acc(NB, 0.0);
        for (int i : bucket[key]) {
            float C[3][4];
            int l[3] = {0, 0, 0};
            int o[3] = {tb.ox, tb.oy, tb.oz}, e[3] = {tb.ex, tb.ey, tb.ez};
            for (int a = 0; a < AXES; ++a) {
                float dist = pos[(size_t)i * AXES + a] * inv;
                int raw = first_node<MAS>(dist);
                foot_weights<MAS>(dist, raw, C[a]);
                l[a] = wrap_node(raw, dims) - o[a];
                assert(l[a] >= 0 && l[a] < e[a]);
            }
            for (int x = 0; x < F; ++x)
                for (int y = 0; y < F; ++y)
                    for (int z = 0; z < (AXES == 3 ? F : 1); ++z) {
                        int idx = ((l[0] + x) * BY + (l[1] + y)) * BZ + l[2] + z;
                        assert(idx >= 0 && idx < NB);
                        acc[idx] += (double)C[0][x] * C[1][y] * (AXES == 3 ? C[2][z] : 1.0f) * W[i];
                    }
        }
        const bool empty = bucket[key].empty();
        for (int c = 0; c < NB; ++c) {
            const int bz = c % BZ, by = (c / BZ) % BY, bx = c / (BZ * BY);
            if (bx >= tb.ex + F - 1 || by >= tb.ey + F - 1 || (AXES == 3 && bz >= tb.ez + F - 1)) {
                assert(acc[c] == 0.0);
                continue;
            }
            const bool inner = bx >= F - 1 && bx < tb.ex && by >= F - 1 && by < tb.ey &&
                               (AXES == 2 || (bz >= F - 1 && bz < tb.ez));
            if (inner) {
                long long node = AXES == 3
                    ? ((long long)(tb.ox + bx) * dims + (tb.oy + by)) * dims + (tb.oz + bz)
                    : (long long)(tb.ox + bx) * dims + (tb.oy + by);
                assert(node >= 0 && node < ncell);
                grid[node] = acc[c];
                writes[node]++;
            } else if (!empty) {
                int h = halo_index<F, AXES>(bx, by, bz, tb.ex, tb.ey, tb.ez);
                assert(h >= 0 && h < (halo_count<F, AXES>(tb.ex, tb.ey, tb.ez)));
                assert(part[key * g.pslot + h] == -1e300);   // each slot written once
                part[key * g.pslot + h] = acc[c];
            }
        }
    }
    if (F > 1) {
        for (long long key = 0; key < g.ntiles; ++key) {
            TileBox tb = tile_box<AXES>((int)key, g);
            const int ex = tb.ex, ey = tb.ey, ez = tb.ez;
            const int nbx = std::min(F - 1, ex), nby = std::min(F - 1, ey), nbz = (AXES == 3) ? std::min(F - 1, ez) : 0;
            const int sa = nbx * ey * ez;
            const int sb = (ex - nbx) * nby * ez;
            const int sc = (AXES == 3) ? (ex - nbx) * (ey - nby) * nbz : 0;
            for (int t = 0; t < sa + sb + sc; ++t) {
                int lx, ly, lz;
                if (t < sa) { lz = t % ez; ly = (t / ez) % ey; lx = t / (ez * ey); }
                else if (t < sa + sb) { int u = t - sa; lz = u % ez; ly = (u / ez) % nby; lx = nbx + u / (ez * nby); }
                else { int u = t - sa - sb; lz = u % nbz; ly = nby + (u / nbz) % (ey - nby); lx = nbx + u / (nbz * (ey - nby)); }
                assert(lx < ex && ly < ey && lz < ez);
                int itx[4], iox[4], iex[4], ity[4], ioy[4], iey[4], itz[4], ioz[4], iez[4];
                int nx = axis_images<F>(tb.tx, g.nt, dims, g.orig, itx, iox, iex);
                int ny = axis_images<F>(tb.ty, g.nt, dims, g.orig, ity, ioy, iey);
                int nz = 1; itz[0] = 0; ioz[0] = 0; iez[0] = 1;
                if (AXES == 3) nz = axis_images<F>(tb.tz, g.ntz, dims, g.orig, itz, ioz, iez);
                assert(nx * ny * nz <= 64);
                double v = 0.0;
                for (int a = 0; a < nx; ++a) {
                    int bx = lx + iox[a];
                    if (bx >= iex[a] + F - 1) continue;
                    for (int b = 0; b < ny; ++b) {
                        int by = ly + ioy[b];
                        if (by >= iey[b] + F - 1) continue;
                        for (int c = 0; c < nz; ++c) {
                            int bz = lz + ioz[c];
                            if (AXES == 3 && bz >= iez[c] + F - 1) continue;
                            long long vkey = ((long long)itx[a] * g.nt + ity[b]) * g.ntz + itz[c];
                            if (bucket[vkey].empty()) continue;
                            int h = halo_index<F, AXES>(bx, by, bz, iex[a], iey[b], iez[c]);
                            assert(h >= 0 && h < (halo_count<F, AXES>(iex[a], iey[b], iez[c])));
                            double pv = part[vkey * g.pslot + h];
                            assert(pv != -1e300);
                            v += pv;
                        }
                    }
                }
                long long node = AXES == 3
                    ? ((long long)(tb.ox + lx) * dims + (tb.oy + ly)) * dims + (tb.oz + lz)
                    : (long long)(tb.ox + lx) * dims + (tb.oy + ly);
                grid[node] = v;
                writes[node]++;
            }
            // uniform-tile finalize (tile_finalize_uniform_kernel): same cells, same sums
            if (AXES == 3 && g.nt >= 2 && dims % kT3 == 0) {
                constexpr int L = F - 1;
                const int NA = L * T * T, NBb = (T - L) * L * T, NC = (T - L) * (T - L) * L;
                assert(NA + NBb + NC == sa + sb + sc);
                const int tz = (int)key % g.nt, ty = ((int)key / g.nt) % g.nt, tx = (int)key / (g.nt * g.nt);
                assert(tx == tb.tx && ty == tb.ty && tz == tb.tz && tb.ox == tx * T && tb.oz == tz * T);
                long long pbase[8];
                for (int i = 0; i < 8; ++i) {
                    const int a = i >> 2, b = (i >> 1) & 1, c = i & 1;
                    const int ix = a ? (tx == 0 ? g.nt - 1 : tx - 1) : tx;
                    const int iy = b ? (ty == 0 ? g.nt - 1 : ty - 1) : ty;
                    const int iz = c ? (tz == 0 ? g.nt - 1 : tz - 1) : tz;
                    const long long vkey = ((long long)ix * g.nt + iy) * g.nt + iz;
                    pbase[i] = bucket[vkey].empty() ? -1 : vkey * g.pslot;
                }
                for (int t = 0; t < NA + NBb + NC; ++t) {
                    int lx, ly, lz;
                    if (t < NA) { lz = t % T; ly = (t / T) % T; lx = t / (T * T); }
                    else if (t < NA + NBb) { int u = t - NA; lz = u % T; ly = (u / T) % L; lx = L + u / (T * L); }
                    else { int u = t - NA - NBb; lz = u % L; ly = L + (u / L) % (T - L); lx = L + u / (L * (T - L)); }
                    double v = 0.0;
                    for (int i = 0; i < 8; ++i) {
                        const int a = i >> 2, b = (i >> 1) & 1, c = i & 1;
                        const bool on = (a == 0 || lx < L) && (b == 0 || ly < L) && (c == 0 || lz < L);
                        if (!(on && pbase[i] >= 0)) continue;
                        const int h = halo_index<F, 3>(lx + a * T, ly + b * T, lz + c * T, T, T, T);
                        assert(h >= 0 && h < g.pslot);
                        const double pv = part[pbase[i] + h];
                        assert(pv != -1e300);
                        v += pv;
                    }
                    const long long node = ((long long)(tb.ox + lx) * dims + (tb.oy + ly)) * dims + (tb.oz + lz);
                    assert(node >= 0 && node < ncell);
                    if (v != grid[node]) { printf("  uniform finalize differs at node %lld\n", node); assert(0); }
                    uni_checked++;
                }
            }
        }
    }
    int bad = 0;
    double tot = 0;
    for (long long c = 0; c < ncell; ++c) tot += std::fabs(ref[c]);
    for (long long c = 0; c < ncell; ++c) {
        if (writes[c] != 1) { if (bad < 3) printf("  node %lld written %d times\n", c, writes[c]); ++bad; }
        if (std::fabs(grid[c] - ref[c]) > 1e-9 * (std::fabs(ref[c]) + tot / ncell)) {
            if (bad < 3) printf("  node %lld got %.17g ref %.17g\n", c, grid[c], ref[c]);
            ++bad;
        }
    }
    printf("MAS %d axes %d dims %3d n %6d pattern %d w %d: %s (nt %d pslot %d)\n", MAS, AXES, dims, n,
           pattern, (int)weighted, bad ? "FAIL" : "ok", g.nt, g.pslot);
    return bad;
}

template <int MAS>
static int all_dims() {
    int bad = 0;
#ifdef PYL_EMUL_QUICK   // the sanitizer build: fewer sizes, every code path
    const int d3[] = {1, 2, 3, 5, 16, 17, 33, 48};
    const int d2[] = {1, 3, 64, 65, 129};
#else
    const int d3[] = {1, 2, 3, 4, 5, 7, 8, 15, 16, 17, 20, 24, 31, 32, 33, 40, 48, 50, 64};
    const int d2[] = {1, 2, 3, 5, 16, 63, 64, 65, 100, 128, 129, 200};
#endif
    unsigned seed = 1;
    for (int d : d3)
        for (int pat = 0; pat < 3; ++pat)
            bad += run_case<MAS, 3>(d, 3000, pat, pat == 2, seed++);
    for (int d : d2)
        for (int pat = 0; pat < 3; ++pat)
            bad += run_case<MAS, 2>(d, 3000, pat, pat == 1, seed++);
    return bad;
}

// Dealing of one wave's round set (deal_hole / deal_record_slot): for uniform,
// skewed, few-class and one-class record sets, full and partial, every record
// gets its own slot, exactly the unused slots are marked kNoRecord, and with no
// excess copies every 16-lane group holds at most one record per class.
template <int KR>
static int check_dealing(unsigned seed) {
    constexpr int WB = KR * 64;
    constexpr unsigned G = 4 * KR;
    std::mt19937 rng(seed);
    std::normal_distribution<float> Nd(0.0f, 1.0f);
    int bad = 0;
    long long sets = 0, excess = 0;
    for (int trial = 0; trial < 4000; ++trial) {
        const int nb = (trial % 5 == 0) ? 1 + (int)(rng() % WB) : WB;
        const int mode = trial % 4;
        std::vector<unsigned> cls(nb);
        for (int i = 0; i < nb; ++i) {
            if (mode == 0) cls[i] = rng() % 16;
            else if (mode == 1) cls[i] = std::min(15u, (unsigned)(std::fabs(Nd(rng)) * 4.0f));
            else if (mode == 2) cls[i] = rng() % 3;
            else cls[i] = 7;
        }
        unsigned cnt[16] = {0};
        std::vector<unsigned> rank(nb);
        for (int i = 0; i < nb; ++i) rank[i] = cnt[cls[i]]++;
        unsigned H[16], X[16], E = 0, holes = 0;
        for (int j = 0; j < 16; ++j) {
            H[j] = holes; X[j] = E;
            holes += cnt[j] < G ? G - cnt[j] : 0u;
            E += cnt[j] > G ? cnt[j] - G : 0u;
        }
        std::vector<unsigned short> si(WB, 0x1234);
        std::vector<unsigned char> ht(WB, 0xee);
        for (unsigned L = 0; L < G; ++L)
            for (unsigned j = 0; j < 16; ++j) deal_hole<KR>(L, j, cnt[j], H[j], E, ht.data(), si.data());
        std::vector<int> owner(WB, -1);
        for (int i = 0; i < nb; ++i) {
            const int sl = deal_record_slot<KR>(cls[i], rank[i], rank[i] < G ? 0u : X[cls[i]], ht.data());
            if (sl < 0 || sl >= WB || owner[sl] != -1 || si[sl] == kNoRecord) { ++bad; continue; }
            owner[sl] = i;
            si[sl] = (unsigned short)i;
        }
        int empty = 0;
        for (int sl = 0; sl < WB; ++sl) {
            if (owner[sl] == -1) { ++empty; if (si[sl] != kNoRecord) ++bad; }
        }
        if (empty != WB - nb) ++bad;
        if (E == 0) {
            for (int gq = 0; gq < WB / 16; ++gq) {
                unsigned seen = 0;
                for (int l = 0; l < 16; ++l) {
                    const int o = owner[gq * 16 + l];
                    if (o < 0) continue;
                    if (seen & (1u << cls[o])) ++bad;
                    seen |= 1u << cls[o];
                }
            }
        }
        ++sets; excess += E;
    }
    printf("dealing KR %d: %lld sets, %lld excess copies: %s\n", KR, sets, excess, bad ? "FAIL" : "ok");
    return bad;
}

// Uniform 3D tiles: box_base_uni (unwrapped first node mod 16) equals the
// general box_base for every record of its home tile, also at the
// periodic edges.
template <int MAS>
static int check_box_uni(unsigned seed) {
    int bad = 0;
    long long n = 0;
    std::mt19937 rng(seed);
    const int dims_list[] = {32, 48, 64, 96, 128, 512};
    for (int dims : dims_list) {
        TileGeom g = make_geom(dims, 3, MAS);
        std::vector<int> orig(g.nt + 1);
        for (int i = 0; i <= g.nt; ++i) orig[i] = tile_origin(i, g.nt, dims);
        g.orig = orig.data();
        if (!(g.nt >= 2 && dims % kT3 == 0)) { ++bad; continue; }
        for (int i = 0; i < 20000; ++i) {
            float d[3];
            for (int a = 0; a < 3; ++a) {
                const int k = (int)(rng() % (dims + 1));
                const float fr[7] = {0.0f, 0.5f, 0.25f, 0.999999f, 0.5000001f, 0.4999999f, 0.75f};
                float x = k + fr[rng() % 7];
                if (rng() % 5 == 0) x = std::nextafter((float)dims, 0.0f);
                if (x >= (float)dims) x = 0.0f;
                d[a] = x;
            }
            const float4 rec = make_float4(d[0], d[1], d[2], 1.0f);
            const int tx = home_tile<MAS, true>(d[0], g), ty = home_tile<MAS, true>(d[1], g),
                      tz = home_tile<MAS, true>(d[2], g);
            const TileBox tb = tile_box<3>((tx * g.nt + ty) * g.ntz + tz, g);
            const int bb = box_base<MAS, 3>(rec, g, tb);
            if (bb != box_base_uni<MAS>(rec)) ++bad;
            ++n;
        }
    }
    printf("uniform box MAS %d: %lld records: %s\n", MAS, n, bad ? "FAIL" : "ok");
    return bad;
}

// Uniform 3D tiles: the deposit's region-by-region halo loops (restated from
// tile_deposit_kernel) visit every halo node of the box once, at its
// halo_index slot, and the interior loop the remaining nodes; the finalize's
// band_cell enumerates the band and its partial slots.
template <int MAS>
static int check_uni_final() {
    constexpr int F = MAS, T = kT3, B = T + F - 1, H = 2 * (F - 1), I = T - (F - 1);
    std::vector<int> seen(B * B * B, 0);
    int bad = 0;
    for (int i = 0; i < I * I * I; ++i) {
        const int x = F - 1 + i / (I * I), y = F - 1 + (i / I) % I, z = F - 1 + i % I;
        ++seen[(x * B + y) * B + z];
    }
    auto halo = [&](int slot, int x, int y, int z) {
        ++seen[(x * B + y) * B + z];
        if (halo_index<F, 3>(x, y, z, T, T, T) != slot) ++bad;
    };
    for (int i = 0; i < H * B * B; ++i) {
        const int xi = i / (B * B), yz = i - xi * (B * B);
        halo(i, xi < F - 1 ? xi : xi + I, yz / B, yz % B);
    }
    constexpr int SA = H * B * B;
    for (int i = 0; i < I * H * B; ++i) {
        const int x = F - 1 + i / (H * B), r = i % (H * B), yi = r / B, z = r - yi * B;
        halo(SA + i, x, yi < F - 1 ? yi : yi + I, z);
    }
    constexpr int SB = SA + I * H * B;
    for (int i = 0; i < I * I * H; ++i) {
        const int x = F - 1 + i / (I * H), r = i % (I * H), y = F - 1 + r / H, zi = r % H;
        halo(SB + i, x, y, zi < F - 1 ? zi : zi + I);
    }
    for (int c = 0; c < B * B * B; ++c) if (seen[c] != 1) ++bad;
    if (SB + I * I * H != halo_count<F, 3>(T, T, T)) ++bad;
    // finalize side: band_cell visits every band cell (some coordinate below
    // F-1) once, and its closed-form slots equal halo_index for every image
    constexpr int L = F - 1;
    const int nband = T * T * T - I * I * I;
    std::vector<int> cell(T * T * T, 0);
    for (int t = 0; t < nband; ++t) {
        int lx, ly, lz, base, dA, dB, dC;
        band_cell<MAS>(t, lx, ly, lz, base, dA, dB, dC);
        if (lx < 0 || lx >= T || ly < 0 || ly >= T || lz < 0 || lz >= T) { ++bad; continue; }
        if (!(lx < L || ly < L || lz < L)) ++bad;
        ++cell[(lx * T + ly) * T + lz];
        for (int i = 0; i < 8; ++i) {
            const int a = i >> 2, b = (i >> 1) & 1, c = i & 1;
            if (!((a == 0 || lx < L) && (b == 0 || ly < L) && (c == 0 || lz < L))) continue;
            if (base + a * dA + b * dB + c * dC != halo_index<F, 3>(lx + a * T, ly + b * T, lz + c * T, T, T, T))
                ++bad;
        }
    }
    for (int c = 0; c < T * T * T; ++c) {
        const int lx = c / (T * T), ly = (c / T) % T, lz = c % T;
        if (cell[c] != ((lx < L || ly < L || lz < L) ? 1 : 0)) ++bad;
    }
    printf("uniform final loops MAS %d: %s\n", MAS, bad ? "FAIL" : "ok");
    return bad;
}

// Deposit order buckets: non-decreasing in the record count (so larger tiles
// never sort after smaller ones), 0 only for empty tiles, within the table.
static int check_count_bucket() {
    int bad = 0, prev = 0;
    for (unsigned long long n = 0; n < (1ull << 22); ++n) {
        const int b = count_bucket(n);
        if (b < prev || b < 0 || b >= kOrderBuckets || (b == 0) != (n == 0)) ++bad;
        prev = b;
    }
    for (int s = 22; s < 64; ++s) {
        const int b = count_bucket(1ull << s);
        if (b < prev || b >= kOrderBuckets) ++bad;
        prev = b;
    }
    printf("count buckets: %s\n", bad ? "FAIL" : "ok");
    return bad;
}

// fix64 (64-bit fixed point of a float count of quanta) against exact arithmetic:
// within one quantum everywhere below 2^62 except at ties (odd multiples of 2^31,
// within 128 quanta = 2^-24 of |s|), truncation toward zero below 2^31, and exact
// for integers that are not ties.
static int check_fix64() {
    std::mt19937_64 rng(29);
    std::uniform_real_distribution<double> ex(-30.0, 62.0);
    int bad = 0;
    auto one = [&](float s) {
        const long long r = (long long)fix64(s);
        const long double err = (long double)r - (long double)s;
        const bool tie = std::fabs(std::fmod((double)s, 4294967296.0)) == 2147483648.0;
        if (!(std::fabs((double)err) < 1.0 || (tie && std::fabs((double)err) <= 128.0))) ++bad;
        if (std::fabs(s) < 2147483648.0f && r != (long long)std::trunc((double)s)) ++bad;
        if (!tie && s == std::trunc(s) && (long double)r != (long double)s) ++bad;
    };
    for (int i = 0; i < 2000000; ++i) {
        float s = (float)std::ldexp(1.0 + (double)(rng() >> 11) * 0x1p-53, 0);
        s = std::ldexp(s, (int)std::floor(ex(rng)));
        if (std::fabs(s) >= 0x1p62f) continue;
        one((rng() & 1) ? s : -s);
    }
    const float edge[] = {0.0f, -0.0f, 0.5f, -0.5f, 1.0f, -1.0f, 0x1p31f, -0x1p31f, 0x1p32f, -0x1p32f,
                          3 * 0x1p31f, -3 * 0x1p31f, 5 * 0x1p31f, -5 * 0x1p31f, 0x1.fffffep61f,
                          -0x1.fffffep61f, 0x1.fffffep30f, -0x1.fffffep30f, 1e-30f, -1e-30f};
    for (float s : edge) one(s);
    printf("fix64: %s\n", bad ? "FAIL" : "ok");
    return bad;
}

// Two-level sort: every tile's bucket is in range; the block window made from any
// tile of a bucket is the bucket's kBkt x kBkt x kBktZ tiles, gives each of its tiles its own slot
// below kBktSlots and maps the slot back to the tile's key; tiles of the
// neighbouring buckets fall outside; the window survives the chunk round trip
// (wbase = a * nt + b, wz = c) the scatter pass decodes.
static int check_block_window() {
    int bad = 0;
    for (int dims : {256, 272, 400, 480, 512, 1000, 1024, 2048}) {
        TileGeom g = make_geom(dims, 3, PYL_TSC);
        g.wmode = kWinBlock;
        const int nbt = bucket_counts_xy(g.nt), nbtz = bucket_counts_z(g.ntz);
        const long long nbk = (long long)nbt * nbt * nbtz;
        if (nbk > kMaxBuckets) continue;          // the host keeps the one-pass sort
        std::vector<std::vector<int3>> tiles(nbk);
        for (int x = 0; x < g.nt; ++x)
            for (int y = 0; y < g.nt; ++y)
                for (int z = 0; z < g.nt; ++z) {
                    const int3 t = make_int3(x, y, z);
                    const int b = tile_bucket(t, nbt, nbtz);
                    if (b < 0 || b >= nbk) { ++bad; continue; }
                    tiles[b].push_back(t);
                }
        for (long long b = 0; b < nbk; ++b) {
            if (tiles[b].empty()) { ++bad; continue; }
            const int3 t0 = tiles[b][tiles[b].size() / 2];
            const Win w = make_win(g, t0, 0, 0);
            if (w.slots != kBktSlots || w.a % kBkt || w.b % kBkt || w.c % kBktZ) ++bad;
            const int wbase = w.a * g.nt + w.b;
            if (wbase / g.nt != w.a || wbase % g.nt != w.b) ++bad;
            std::vector<char> used(kBktSlots, 0);
            for (const int3 &t : tiles[b]) {
                const unsigned key = tiles_key(t, g, 0);
                const unsigned sl = win_slot(t, key, w, g, 0);
                if (sl >= (unsigned)kBktSlots || used[sl]) { ++bad; continue; }
                used[sl] = 1;
                if (win_key((int)sl, w, g, 0) != (long long)key) ++bad;
            }
            // tiles one step outside the block along each axis (when they exist)
            for (int a = 0; a < 3; ++a)
                for (int side : {-1, a == 2 ? kBktZ : kBkt}) {
                    int3 t = make_int3(w.a, w.b, w.c);
                    int *c = a == 0 ? &t.x : a == 1 ? &t.y : &t.z;
                    *c += side;
                    if (*c < 0 || *c >= g.nt) continue;
                    if (win_slot(t, tiles_key(t, g, 0), w, g, 0) < (unsigned)kBktSlots) ++bad;
                }
        }
    }
    printf("block window: %s\n", bad ? "FAIL" : "ok");
    return bad;
}

int main() {
    int bad = check_dealing<4>(11) + check_dealing<2>(12) + check_count_bucket() + check_fix64();
    bad += check_block_window();
    bad += check_uni_final<PYL_NGP>() + check_uni_final<PYL_CIC>() + check_uni_final<PYL_TSC>() +
           check_uni_final<PYL_PCS>();
    bad += check_box_uni<PYL_CIC>(21) + check_box_uni<PYL_TSC>(22) + check_box_uni<PYL_PCS>(23);
    bad += all_dims<PYL_NGP>() + all_dims<PYL_CIC>() + all_dims<PYL_TSC>() + all_dims<PYL_PCS>();
    printf("uniform finalize cells checked: %lld\n", uni_checked);
    printf(bad ? "FAILURES: %d\n" : "all ok\n", bad);
    return bad != 0;
}
