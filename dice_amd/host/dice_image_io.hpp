// Image files for the host mirror: the reference's .rawi format
// (src/rawi/DICe_Rawi.cpp) and baseline TIFF (utils::read_image,
// src/ioutils/DICe_ImageIO.cpp:353-520, which decodes TIFF through OpenCV's
// IMREAD_GRAYSCALE).  Errors throw std::runtime_error (the reference prints
// and exits for .rawi, throws for other files).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace dice_hip {

// Image_File_Type of DICe.h:695-706.  This build reads RAWI and TIFF files;
// the readers throw std::runtime_error for the other types.
enum Image_File_Type { RAWI = 0, TIFF, JPEG, PNG, NETCDF, CINE, BMP, VIDEO, MAX_IMAGE_FILE_TYPE, NO_SUCH_IMAGE_FILE_TYPE };

// utils::image_file_type (DICe_ImageIO.cpp:134-188): by substring of the file
// name, in the reference's order (".rawi", ".cine", "tif"/"tiff"/"TIF"/"TIFF",
// jpg, bmp, png, ".nc", ".avi"/".mp4").
Image_File_Type image_file_type(const std::string &file_name);

// utils::read_image_dimensions (DICe_ImageIO.cpp:190-260).
void read_image_dimensions(const std::string &file_name, int &width, int &height);

// utils::read_image (DICe_ImageIO.cpp:353-520) with layout_right: row-major
// intensities of the region [sub_offset_x, +sub_w) x [sub_offset_y, +sub_h)
// (sub_w = 0 / sub_h = 0: to the right / bottom edge of the image).
// A region of a .rawi file throws, as in the reference (DICe_ImageIO.cpp:286-290);
// .rawi files must hold 4-byte values (the reference's float build rejects
// other sizes, DICe_Rawi.cpp:79-83).  TIFF: one 8- or 16-bit grey sample
// per pixel (BlackIsZero or WhiteIsZero) or an 8-bit palette index, strips,
// no compression, LZW (with or without the horizontal predictor) or
// PackBits, either byte order.  Intensities are 8-bit values, as OpenCV's
// IMREAD_GRAYSCALE delivers them: 16-bit samples scale to
// (v * 255 + 32767) / 65535, WhiteIsZero samples are inverted, palette
// entries become grey by the BT.601 weights (a grey palette, R = G = B,
// gives its own level).
void read_image(const std::string &file_name, std::vector<float> &intensities, int &width, int &height,
                int sub_offset_x = 0, int sub_offset_y = 0, int sub_w = 0, int sub_h = 0);

// The samples of a TIFF file as stored (8- or 16-bit, WhiteIsZero not
// inverted), for the 16-bit storage path of the library; no counterpart in
// the reference, whose reader reduces every TIFF to 8 bits.  bits is set to
// the file's bits per sample.
void read_tiff_samples(const std::string &file_name, std::vector<uint16_t> &samples, int &width, int &height,
                       int &bits);

// utils::read_rawi_image / write_rawi_image (DICe_Rawi.cpp:64-140),
// layout_right: header uint32 width, height, bytes per value, then the
// values row by row in the machine's byte order.  read_rawi_image throws if
// the file's value size differs from sizeof(T).
template <typename T>
void read_rawi_image(const std::string &file_name, std::vector<T> &values, int &width, int &height);
template <typename T>
void write_rawi_image(const std::string &file_name, int width, int height, const T *values);

}  // namespace dice_hip
