// svd_integrated_imager_gpu — the integrated CLI (SVDIntegratedImager.cu:1763-1854) on
// libsvdimager:
//   svd_integrated_imager_gpu MeasurementSet_columns Image_Size Cell_Size Snapshot_Index Output.fits
// The first argument is a directory of the MeasurementSet's columns exported as .npy files (see
// svdi_ms_read_snapshot); the rest, the messages, the stdout line and the exit codes (2 for a
// wrong argument count, 1 for any failure) are the reference's.  The snapshot is imaged with
// the library's default options (the MFMA gridder, integrated semantics, the dirty image) and
// the Chebyshev report printed as the reference's FIpipeDevice does.
#include <chrono>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "svdimager.h"

namespace {

void print_usage(const char* program) {
    std::cerr << "Usage: " << program
              << " MeasurementSet.ms Image_Size Cell_Size Snapshot_Index Output_Name.fits\n";
}

// Frees the snapshot and the context on every path out of main.
struct Resources {
    svdi_ms_snapshot snap{};
    svdi_context* ctx = nullptr;
    ~Resources() {
        svdi_ms_free_snapshot(&snap);
        if (ctx) svdi_destroy(ctx);
    }
};

}  // namespace

int main(int argc, char** argv) {
    if (argc != 6) {
        print_usage(argv[0]);
        return 2;
    }
    try {
        const std::string ms_path(argv[1]);
        const size_t image_size = static_cast<size_t>(std::stoul(argv[2]));
        const float cell_size = std::stof(argv[3]);
        const size_t snapshot_index = static_cast<size_t>(std::stoul(argv[4]));
        const std::string output_name(argv[5]);
        if (image_size == 0) throw std::runtime_error("Image_Size must be greater than zero");
        if (cell_size <= 0.0f) throw std::runtime_error("Cell_Size must be greater than zero");

        const auto total_start = std::chrono::high_resolution_clock::now();
        std::cerr << "Reading Measurement Set snapshots by TIME: " << ms_path << '\n';
        Resources res;
        if (svdi_ms_read_snapshot(ms_path.c_str(), snapshot_index, &res.snap) != SVDI_OK)
            throw std::runtime_error(svdi_last_error());
        std::cerr << "Found " << res.snap.total_snapshots << " TIME snapshot(s)\n";
        const svdi_snapshot& data = res.snap.data;
        const size_t num_samples = data.num_rows * data.num_channels;
        std::cerr << "Selected snapshot " << res.snap.index << " time " << res.snap.time << " rows "
                  << data.num_rows << " samples " << num_samples << '\n';
        std::cerr << "Preprocessing " << num_samples << " visibility sample(s) on GPU\n";
        if (svdi_create(&res.ctx, 0) != SVDI_OK) throw std::runtime_error(svdi_last_error());
        std::vector<float> dirty_image(image_size * image_size, 0.0f);
        std::cerr << "Running imager on GPU preprocess buffers\n";
        svdi_options opts;
        svdi_default_options(&opts);
        opts.verbose = 1;
        if (svdi_image_host_snapshot(res.ctx, &data, dirty_image.data(), image_size, cell_size, &opts, nullptr) !=
            SVDI_OK)
            throw std::runtime_error(std::string("FIpipeDevice failed: ") + svdi_last_error());
        if (svdi_fits_write(output_name.c_str(), dirty_image.data(), static_cast<long>(image_size),
                            static_cast<long>(image_size)) != SVDI_OK)
            throw std::runtime_error("cannot write " + output_name);
        const auto total_stop = std::chrono::high_resolution_clock::now();
        const double total_seconds = std::chrono::duration<double>(total_stop - total_start).count();
        std::cerr << "Wrote image FITS: " << output_name << '\n';
        std::cout << "Total pipeline time: " << total_seconds * 1000.0 << " ms" << std::endl;
        return 0;
    } catch (const std::exception& ex) {
        std::cerr << "Integrated SVD imager failed: " << ex.what() << '\n';
        return 1;
    }
}
