// mkde_cli.cpp -- mgpukde_tiled: the reference's command line (tiled:262-357,
// README.md:19-23) on top of libmkde.
//
//   mgpukde_tiled 0 #points cellsize_ec cellsize_density h_option edge_corr sample_weight
//                 skip_cpu skip_gpu num_cpu_threads num_gpu den_cpu.tif den_gpu.tif serialized
//   mgpukde_tiled 1 points.csv mask_ec.tif mask_density.tif h_option edge_corr sample_weight
//                 skip_cpu skip_gpu num_cpu_threads num_gpu den_cpu.tif den_gpu.tif serialized
//
// h_option: -1 bandwidths from the CSV (mode 1 only), 0 rule of thumb, 1 MLE
// fixed, 2 adaptive (hooke_jeeves then Abramson).  Outputs: the GPU density
// GeoTIFF (BigTIFF, LZW, float32) and pntsCUDA.csv (x, y, h, edge-correction
// weight).  The reference's OpenMP CPU path is not part of this build: with
// skip_cpu = 0 a note is printed and den_cpu.tif is not written.  One GPU is
// used whatever num_gpu asks for.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <string>
#include <vector>

#include "mkde.h"

using Clock = std::chrono::high_resolution_clock;

static double ms_since(Clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

static void usage()
{
    printf("Incorrect arguments provided. Exiting...\n");
    printf("Run in mode 0:\n mgpukde_tiled 0 #points cellsize_edge_correction cellsize_density h_option "
           "enable_edge_corection enable_sample_weight skip_omp_parallel skip_gpu_parallel num_cpu_threads num_gpu "
           "denfn_omp denfn_cuda serialized_mode\n");
    printf("Run in mode 1:\n mgpukde_tiled 1 points_file mask_file_edge_correction mask_file_density h_option "
           "enable_edge_corection enable_sample_weight skip_omp_parallel skip_gpu_parallel num_cpu_threads num_gpu "
           "denfn_omp denfn_cuda serialized_mode\n");
}

static int die(const char *what)
{
    printf("***ERROR in %s: %s\n", what, mkde_last_error());
    return 2;
}

// Mode-0 points: srand(100); x = rand()*100.0f/RAND_MAX; y likewise
// (tiled:407, 1611-1618), with the Microsoft C runtime rand() the reference
// was built against: state = state*214013 + 2531011, output bits 16..30.
static void mode0_points(int64_t n, std::vector<float> &x, std::vector<float> &y)
{
    uint32_t s = 100u;
    x.resize(n);
    y.resize(n);
    for (int64_t i = 0; i < n; ++i) {
        s = s * 214013u + 2531011u;
        const int r1 = (int)((s >> 16) & 0x7fff);
        s = s * 214013u + 2531011u;
        const int r2 = (int)((s >> 16) & 0x7fff);
        x[i] = (float)r1 * 100.0f / (float)32767;
        y[i] = (float)r2 * 100.0f / (float)32767;
    }
}

// The reference's GPU tile size (tiled:1184-1231): whole raster when it fits
// in 20% of the free device memory at 12 bytes per cell, else multiples of
// the GeoTIFF block size chosen by its loop.
static void tile_size(int64_t free_bytes, int bxs, int bys, int64_t ncols, int64_t nrows, int64_t &tx, int64_t &ty)
{
    const double nbytes_block = 4.0 * bxs * bys * 3.0;
    const int64_t nblocks = (int64_t)std::floor((double)free_bytes * 0.2 / nbytes_block);
    const int64_t nbx = (int64_t)std::ceil(ncols * 1.0 / bxs), nby = (int64_t)std::ceil(nrows * 1.0 / bys);
    tx = ncols;
    ty = nrows;
    if (nblocks < nbx * nby) {
        int64_t tby = 1, tbx = 1;
        while (tby < nby / 2) {
            if (tby * tbx < nblocks) {
                tby += 1;
                while (tbx <= nbx / 2) {
                    if (tby * tbx < nblocks) {
                        tbx += 1;
                        break;
                    } else {
                        tbx -= 1;
                        break;
                    }
                }
            } else {
                tby -= 1;
                break;
            }
        }
        tx = tbx * bxs;
        ty = tby * bys;
    }
}

int main(int argc, char **argv)
{
    const auto T1 = Clock::now();
    if (argc != 15) {
        usage();
        return 1;
    }
    const int mode = atoi(argv[1]);
    if (mode != 0 && mode != 1) {
        usage();
        return 1;
    }
    const int hoption = atoi(argv[5]);
    const int update_weights = atoi(argv[6]);
    const int sample_weights = atoi(argv[7]);
    const int skip_cpu = atoi(argv[8]);
    const int skip_gpu = atoi(argv[9]);
    const int ngpu = atoi(argv[11]);
    const char *den_gpu_fn = argv[13];
    if (mode == 0 && hoption == -1) {
        printf("***Error - should never use bandwidth option -1 in mode 0 (i.e., randomly generately points). "
               "Exiting...");
        return 1;
    }
    if (hoption < -1 || hoption > 2) {
        printf("***Error - bandwidth option must be -1, 0, 1 or 2\n");
        return 1;
    }
    if (atoi(argv[14]) != 0) printf("===Serialized raster mode is not used by this build (dense rasters only)\n");
    if (ngpu != 1) printf("===%d GPU(s) requested; this build uses one GPU\n", ngpu);

    // ------------------------------------------------------------ inputs --
    std::vector<float> x, y, w, hs;
    mkde_grid ecg;
    std::vector<float> ecmask;
    int64_t dcols, drows;
    double dgt[6];
    float dnodata = -9999.0f;
    int bxs = 128, bys = 128;
    const char *den_mask_fn = nullptr;
    auto t1 = Clock::now();
    if (mode == 0) {
        const int64_t npts = atoll(argv[2]);
        const float cs_ec = (float)atof(argv[3]);
        const float cs_den = (float)atof(argv[4]);
        if (npts <= 0 || !(cs_ec > 0.0f) || !(cs_den > 0.0f)) {
            usage();
            return 1;
        }
        mode0_points(npts, x, y);
        w.assign(npts, 1.0f);
        const int64_t n_ec = (int64_t)(100 / cs_ec);   // int(100 / CELLSIZE), tiled:429
        ecg = mkde_grid{n_ec, n_ec, 0.0f, 0.0f, cs_ec, -9999.0f};
        ecmask.assign((size_t)n_ec * n_ec, 0.0f);
        dcols = drows = (int64_t)std::ceil(100.0 / (double)cs_den);   // tiled:604-605
        const double g[6] = {0.0, (double)cs_den, 0.0, 100.0, 0.0, -1.0 * (double)cs_den};
        std::copy(g, g + 6, dgt);
    } else {
        const char *pts_fn = argv[2], *ec_fn = argv[3];
        den_mask_fn = argv[4];
        int64_t n = 0;
        if (mkde_csv_read_points(pts_fn, hoption == -1, sample_weights, &n, nullptr, nullptr, nullptr, nullptr))
            return die("ReadSamplePoints");
        x.resize(n);
        y.resize(n);
        w.resize(n);
        hs.resize(n);
        if (mkde_csv_read_points(pts_fn, hoption == -1, sample_weights, &n, x.data(), y.data(), hs.data(), w.data()))
            return die("ReadSamplePoints");
        int64_t nc, nr;
        double gt[6], nd;
        int hasnd;
        if (mkde_tiff_info(ec_fn, &nc, &nr, gt, &nd, &hasnd, nullptr, nullptr)) return die("ReadGeoTIFFRaster");
        ecmask.resize((size_t)nc * nr);
        if (mkde_tiff_read(ec_fn, 0, 0, nc, nr, ecmask.data())) return die("ReadGeoTIFFRaster");
        // ReadGeoTIFFRaster (tiled:2069-2077)
        ecg = mkde_grid{nc, nr, (float)gt[0], (float)(gt[3] + nr * gt[5]), (float)gt[1], (float)nd};
        if (mkde_tiff_info(den_mask_fn, &dcols, &drows, dgt, &nd, &hasnd, &bxs, &bys))
            return die("GeoTIFFReader");
        dnodata = (float)nd;
    }
    printf("...reading in data took %.0f ms\n", ms_since(t1));
    const int64_t npts = (int64_t)x.size();
    printf("===Number of points: %lld\n", (long long)npts);
    printf("===Cell size (edge correction): %f\n", ecg.cellsize);
    printf("===Number of cells (edge correction): %lld\n", (long long)(ecg.ncols * ecg.nrows));
    printf("===Number of cells (density estimation): %lld\n", (long long)(dcols * drows));
    printf("===Bandwidth option: %d\n===Enable edge correction: %d\n===Enable sample weight: %d\n", hoption,
           update_weights, sample_weights);

    mkde_ctx *ctx = nullptr;
    if (mkde_ctx_create(0, &ctx)) return die("mkde_ctx_create");
    if (mkde_points_upload(ctx, npts, x.data(), y.data(), w.data())) return die("points upload");
    float h0 = 0.0f;
    if (mkde_rule_of_thumb(ctx, &h0)) return die("BandWidth2");
    printf("===Rule of thumb bandwidth h0: %.5f\n", h0);

    if (skip_cpu == 0)
        printf("===The OpenMP CPU path is not part of this build (GPU path only); %s is not written\n", argv[12]);

    int rc = 0;
    if (skip_gpu == 0) {
        const auto t1_gpu = Clock::now();
        if (mkde_mask_upload(ctx, &ecg, ecmask.data())) return die("mask upload");
        if (hoption == -1 && mkde_set_bandwidths(ctx, hs.data())) return die("bandwidths");
        int64_t free_b = 0;
        mkde_device_memory(ctx, &free_b, nullptr);
        int64_t TX, TY;
        tile_size(free_b, bxs, bys, dcols, drows, TX, TY);
        printf("...TILE_XSIZE=%lld, TILE_YSIZE=%lld\n", (long long)TX, (long long)TY);
        mkde_raster_writer *wr = nullptr;
        if (mkde_writer_open(den_gpu_fn, dcols, drows, dgt, dnodata, den_mask_fn, &wr)) return die("GeoTIFFWriter");
        mkde_job_opts o{hoption, update_weights, 1, -0.5f, 1};
        mkde_job_stats st{};
        bool first = true;
        int cnt = 1;
        const int64_t ntiles = ((dcols + TX - 1) / TX) * ((drows + TY - 1) / TY);
        std::vector<float> tile;
        for (int64_t yoff = 0; yoff < drows; yoff += TY) {
            for (int64_t xoff = 0; xoff < dcols; xoff += TX) {
                const auto t1t = Clock::now();
                const int64_t tx = std::min(TX, dcols - xoff), ty = std::min(TY, drows - yoff);
                printf("...working on tile %d / %lld\n", cnt, (long long)ntiles);
                mkde_grid tg;
                tile.assign((size_t)tx * ty, 0.0f);
                if (mode == 0) {
                    // AllocateAsciiRaster for a mode-0 tile (tiled:1398): float origin arithmetic
                    const float cs = (float)dgt[1];
                    tg = mkde_grid{tx, ty, (float)xoff * cs, 100.0f - (float)(yoff + ty) * cs, cs, -9999.0f};
                    if (mkde_tile_upload(ctx, &tg, nullptr)) return die("tile upload");
                } else {
                    if (mkde_tiff_read(den_mask_fn, xoff, yoff, tx, ty, tile.data())) return die("tile read");
                    // AsciiRasterFromGeoTIFFTile with the tile geotransform (GeoTIFF.cpp:188-196)
                    const double g0 = dgt[0] + dgt[1] * xoff, g3 = dgt[3] + dgt[5] * yoff;
                    tg = mkde_grid{tx, ty, (float)g0, (float)(g3 + ty * dgt[5]), (float)dgt[1], dnodata};
                    if (mkde_tile_upload(ctx, &tg, tile.data())) return die("tile upload");
                }
                if (first) {
                    if (mkde_run_job(ctx, &o, &st)) return die("KDE job");
                    first = false;
                } else if (mkde_density_tile_resident(ctx)) {
                    return die("KernelDesityEstimation");
                }
                if (mkde_tile_download(ctx, tile.data())) return die("tile download");
                printf("...KernelDesityEstimation on tile %d took %.0f ms\n", cnt, ms_since(t1t));
                if (mkde_writer_put(wr, xoff, yoff, tx, ty, tile.data())) return die("WriteGeoTIFF_NextTile");
                ++cnt;
            }
        }
        if (mkde_writer_close(wr, 1)) return die("GeoTIFFWriter");
        if (hoption == 1) printf("cross validated optimal fixed bandwidth hopt: %.5f\n", st.h_opt);
        if (hoption == 2) printf("h0: %.5f alpha0: %.5f\n", st.h_opt, st.alpha_opt);
        // pntsCUDA.csv: x, y, h, edge-correction weight (tiled:1524), points in
        // the order of their boundary distance when edge correction is on
        std::vector<float> hh(npts), ew(npts);
        if (mkde_get_bandwidths(ctx, hh.data()) || mkde_get_ec_weights(ctx, ew.data())) return die("results");
        std::vector<int64_t> ord(npts);
        std::iota(ord.begin(), ord.end(), 0);
        if (update_weights) {
            std::vector<double> d(npts);
            if (mkde_dist_to_boundary(ctx, d.data())) return die("CalcDist2Boundary");
            std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return d[a] < d[b]; });
        }
        std::vector<float> ox(npts), oy(npts), oh(npts), ow(npts);
        for (int64_t i = 0; i < npts; ++i) {
            ox[i] = x[ord[i]];
            oy[i] = y[ord[i]];
            oh[i] = hh[ord[i]];
            ow[i] = ew[ord[i]];
        }
        if (mkde_csv_write_points("pntsCUDA.csv", npts, ox.data(), oy.data(), oh.data(), ow.data()))
            return die("WriteSamplePoints");
        printf("\n>>>>>>Computation on GPU took %.0f ms in total\n\n", ms_since(t1_gpu));
    }
    mkde_ctx_destroy(ctx);
    printf("Done...\n\n+++In total it took %.0f ms\n", ms_since(T1));
    return rc;
}
