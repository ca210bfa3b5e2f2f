// sharedlibrary_gpu — the archive CLI (archive/pipe_interface.cpp:139-186) on libsvdimager:
//   sharedlibrary_gpu Visreal.fits Visimag.fits B.fits V.fits Image_Size Number_of_Rows Cell_Size Output.fits
// Reads the four FITS inputs, calls FIpipe, writes the dirty image as a FITS file.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "svdimager.h"

namespace {

// read_fits_image (archive/pipe_interface.cpp:12-96), including its DEBUG lines.
float* read_fits_image(const char* filename, long* naxes) {
    float* data = nullptr;
    const int rc = svdi_fits_read(filename, &data, naxes);
    if (rc != SVDI_OK) {
        std::fprintf(stderr, "ERROR: could not read FITS image %s\n", filename);
        return nullptr;
    }
    const long total = naxes[0] * naxes[1];
    const int naxis = naxes[1] == 1 ? 1 : 2;
    std::printf("DEBUG: Reading %s, naxis=%d\n", filename, naxis);
    if (naxis == 1) {
        std::printf("DEBUG: 1D array, naxes[0]=%ld, total_elements=%ld\n", naxes[0], total);
    } else {
        std::printf("DEBUG: 2D array, naxes[0]=%ld, naxes[1]=%ld, total_elements=%ld\n", naxes[0], naxes[1], total);
    }
    std::printf("DEBUG: Allocated %.2f MB for %ld elements\n", (total * sizeof(float)) / (1024.0 * 1024.0), total);
    std::printf("DEBUG: Successfully read %ld elements, first value=%.6e, last value=%.6e\n", total,
                data[0], data[total - 1]);
    return data;
}

}  // namespace

int main(int argc, char* argv[]) {
    if (argc < 9) {
        std::fprintf(stderr,
                     "Usage: %s Visreal.fits Visimag.fits B.fits V.fits Image_Size Number_of_Rows "
                     "Cell_Size Output.fits\n",
                     argv[0]);
        return 2;
    }
    long imasize[2];
    float* visreal = read_fits_image(argv[1], imasize);
    if (!visreal) return 1;
    const long n_visreal = imasize[0] * imasize[1];
    float* visimag = read_fits_image(argv[2], imasize);
    if (!visimag) return 1;
    const long n_visimag = imasize[0] * imasize[1];
    float* bin = read_fits_image(argv[3], imasize);
    if (!bin) return 1;
    const long n_bin = imasize[0] * imasize[1];
    float* vin = read_fits_image(argv[4], imasize);
    if (!vin) return 1;
    const long n_vin = imasize[0] * imasize[1];
    size_t image_size, num_baselines;
    float cell_size;
    try {
        image_size = std::stoul(argv[5]);
        num_baselines = std::stoul(argv[6]);
        cell_size = std::stof(argv[7]);
    } catch (const std::exception&) {
        std::fprintf(stderr, "Invalid Image_Size / Number_of_Rows / Cell_Size\n");
        return 2;
    }
    // FIpipe reads Number_of_Rows samples of Visreal / Visimag, 3 x Number_of_Rows of B and 9 of V
    // (the reference copies them without checking, archive/pipe_interface.cpp:176-178).
    if (num_baselines > static_cast<size_t>(n_visreal) || num_baselines > static_cast<size_t>(n_visimag) ||
        3 * num_baselines > static_cast<size_t>(n_bin) || n_vin < 9) {
        std::fprintf(stderr,
                     "Number_of_Rows %zu does not fit the inputs (Visreal %ld, Visimag %ld, B %ld, V %ld values)\n",
                     num_baselines, n_visreal, n_visimag, n_bin, n_vin);
        return 1;
    }
    float* dirty = static_cast<float*>(std::malloc(image_size * image_size * sizeof(float)));
    if (!dirty) return 1;
    if (FIpipe(visreal, visimag, bin, vin, dirty, num_baselines, image_size, cell_size) != 0) {
        std::free(dirty);
        return 1;                                          // FIpipe has printed the reason
    }
    const int status = svdi_fits_write(argv[8], dirty, static_cast<long>(image_size), static_cast<long>(image_size));
    if (status != SVDI_OK) {
        std::fprintf(stderr, "Error writing FITS image\n");
        return 1;
    }
    svdi_free(visreal);
    svdi_free(visimag);
    svdi_free(bin);
    svdi_free(vin);
    std::free(dirty);
    return 0;
}
