// Image files for the host mirror: .rawi (DICe_Rawi.cpp) and baseline TIFF
// (the files utils::read_image, DICe_ImageIO.cpp:353-520, reads through
// OpenCV).  The TIFF decoder follows the TIFF 6.0 specification: the image
// file header and first IFD, strips, compression 1 (none), 5 (LZW, MSB-first
// codes of 9 to 12 bits with the early width change) and 32773 (PackBits),
// predictor 2 (horizontal differencing).
#include "dice_image_io.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>

namespace dice_hip {

namespace {

[[noreturn]] void fail(const std::string &file, const std::string &what) {
  throw std::runtime_error("image file " + file + ": " + what);
}

std::vector<unsigned char> slurp(const std::string &file) {
  std::ifstream in(file, std::ios::binary);
  if (!in) fail(file, "cannot open");
  return std::vector<unsigned char>((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
}

struct Tiff {
  int width = 0, height = 0, bits = 8, compression = 1, photometric = 1, predictor = 1;
  int samples_per_pixel = 1, planar = 1, rows_per_strip = 0;
  std::vector<uint32_t> strip_offsets, strip_counts;
  std::vector<uint16_t> colormap;  // palette images: 3 * 2^bits values (red, green, blue)
  bool big_endian = false;
};

struct Reader {
  const std::vector<unsigned char> &b;
  const std::string &file;
  bool be;
  uint32_t u16(size_t o) const {
    if (o + 2 > b.size()) fail(file, "truncated");
    return be ? (b[o] << 8 | b[o + 1]) : (b[o] | b[o + 1] << 8);
  }
  uint32_t u32(size_t o) const {
    if (o + 4 > b.size()) fail(file, "truncated");
    return be ? ((uint32_t)b[o] << 24 | b[o + 1] << 16 | b[o + 2] << 8 | b[o + 3])
              : (b[o] | b[o + 1] << 8 | b[o + 2] << 16 | (uint32_t)b[o + 3] << 24);
  }
};

// TIFF 6.0 section 2: header, first IFD; field types SHORT (3) and LONG (4)
// for the tags a strip image needs.
Tiff parse_tiff(const std::vector<unsigned char> &b, const std::string &file) {
  if (b.size() < 8) fail(file, "not a TIFF file");
  Tiff t;
  if (b[0] == 'I' && b[1] == 'I') t.big_endian = false;
  else if (b[0] == 'M' && b[1] == 'M') t.big_endian = true;
  else fail(file, "not a TIFF file");
  const Reader r{b, file, t.big_endian};
  if (r.u16(2) != 42) fail(file, "not a TIFF file");
  const size_t ifd = r.u32(4);
  const uint32_t n = r.u16(ifd);
  bool tiled = false;
  for (uint32_t k = 0; k < n; ++k) {
    const size_t e = ifd + 2 + 12 * (size_t)k;
    const uint32_t tag = r.u16(e), type = r.u16(e + 2), count = r.u32(e + 4);
    const size_t size = type == 3 ? 2 : type == 4 ? 4 : 0;
    auto value = [&](uint32_t i) -> uint32_t {
      if (!size) fail(file, "tag " + std::to_string(tag) + " has an unexpected field type");
      const size_t at = (size_t)count * size <= 4 ? e + 8 + i * size : r.u32(e + 8) + i * size;
      return size == 2 ? r.u16(at) : r.u32(at);
    };
    switch (tag) {
      case 256: t.width = (int)value(0); break;
      case 257: t.height = (int)value(0); break;
      case 258:
        t.bits = (int)value(0);
        for (uint32_t i = 1; i < count; ++i)
          if ((int)value(i) != t.bits) fail(file, "mixed bits per sample");
        break;
      case 259: t.compression = (int)value(0); break;
      case 262: t.photometric = (int)value(0); break;
      case 273:
        t.strip_offsets.resize(count);
        for (uint32_t i = 0; i < count; ++i) t.strip_offsets[i] = value(i);
        break;
      case 277: t.samples_per_pixel = (int)value(0); break;
      case 278: t.rows_per_strip = (int)value(0); break;
      case 279:
        t.strip_counts.resize(count);
        for (uint32_t i = 0; i < count; ++i) t.strip_counts[i] = value(i);
        break;
      case 284: t.planar = (int)value(0); break;
      case 317: t.predictor = (int)value(0); break;
      case 320:
        t.colormap.resize(count);
        for (uint32_t i = 0; i < count; ++i) t.colormap[i] = (uint16_t)value(i);
        break;
      case 322: case 323: case 324: case 325: tiled = true; break;
      default: break;
    }
  }
  if (tiled) fail(file, "tiled TIFF is not supported");
  if (t.width < 1 || t.height < 1) fail(file, "bad dimensions");
  if (t.samples_per_pixel != 1) fail(file, "only one grey sample per pixel is supported");
  if (t.bits != 8 && t.bits != 16) fail(file, "only 8- and 16-bit samples are supported");
  if (t.photometric != 0 && t.photometric != 1 && t.photometric != 3)
    fail(file, "only grey (WhiteIsZero / BlackIsZero) and palette images are supported");
  if (t.photometric == 3 && (t.bits != 8 || t.colormap.size() != 3 * 256u)) fail(file, "bad palette");
  if (t.compression != 1 && t.compression != 5 && t.compression != 32773)
    fail(file, "compression " + std::to_string(t.compression) + " is not supported");
  if (t.predictor != 1 && t.predictor != 2) fail(file, "predictor " + std::to_string(t.predictor) + " is not supported");
  if (t.strip_offsets.empty() || t.strip_offsets.size() != t.strip_counts.size()) fail(file, "bad strips");
  if (t.rows_per_strip <= 0 || t.rows_per_strip > t.height) t.rows_per_strip = t.height;
  return t;
}

// LZW (TIFF 6.0 section 13): codes 256 = Clear, 257 = EndOfInformation,
// the first free code 258; the width grows one code early (at 511, 1023, 2047).
void lzw_decode(const unsigned char *in, size_t n, std::vector<unsigned char> &out, size_t want, const std::string &file) {
  std::vector<uint16_t> prefix(4096);
  std::vector<unsigned char> suffix(4096), first(4096);
  std::vector<uint16_t> length(4096);
  for (int i = 0; i < 256; ++i) { prefix[i] = 0xffff; suffix[i] = (unsigned char)i; first[i] = (unsigned char)i; length[i] = 1; }
  std::vector<unsigned char> stack(4096);
  size_t bitpos = 0;
  int width = 9, next = 258, old = -1;
  const size_t start = out.size();
  auto emit = [&](int code) {
    const int len = length[code];
    int c = code;
    for (int i = len - 1; i >= 0; --i) { stack[i] = suffix[c]; c = prefix[c]; }
    out.insert(out.end(), stack.begin(), stack.begin() + len);
  };
  while (out.size() - start < want) {
    if (bitpos + width > n * 8) break;
    int code = 0;
    for (int i = 0; i < width; ++i, ++bitpos) code = code << 1 | ((in[bitpos >> 3] >> (7 - (bitpos & 7))) & 1);
    if (code == 257) break;
    if (code == 256) {
      width = 9;
      next = 258;
      old = -1;
      continue;
    }
    if (old < 0) {
      if (code > 255) fail(file, "corrupt LZW data");
      emit(code);
      old = code;
      continue;
    }
    if (code < next) {
      emit(code);
      if (next < 4096) {
        prefix[next] = (uint16_t)old; suffix[next] = first[code]; first[next] = first[old];
        length[next] = (uint16_t)(length[old] + 1);
        ++next;
      }
    } else if (code == next && next < 4096) {
      prefix[next] = (uint16_t)old; suffix[next] = first[old]; first[next] = first[old];
      length[next] = (uint16_t)(length[old] + 1);
      ++next;
      emit(code);
    } else {
      fail(file, "corrupt LZW data");
    }
    old = code;
    if (next + 1 >= (1 << width) && width < 12) ++width;
  }
}

// PackBits (TIFF 6.0 section 9).
void packbits_decode(const unsigned char *in, size_t n, std::vector<unsigned char> &out, size_t want) {
  const size_t start = out.size();
  size_t i = 0;
  while (i < n && out.size() - start < want) {
    const int c = (signed char)in[i++];
    if (c >= 0) {
      const size_t len = std::min((size_t)c + 1, n - i);
      out.insert(out.end(), in + i, in + i + len);
      i += len;
    } else if (c != -128 && i < n) {
      out.insert(out.end(), (size_t)(1 - c), in[i++]);
    }
  }
}

// The grey samples of the whole image, row-major, as stored.
std::vector<uint16_t> tiff_samples(const std::string &file, int &width, int &height, int &bits, int &photometric,
                                   std::vector<uint16_t> *colormap = nullptr) {
  const std::vector<unsigned char> b = slurp(file);
  const Tiff t = parse_tiff(b, file);
  const size_t bps = t.bits / 8, row_bytes = (size_t)t.width * bps;
  std::vector<unsigned char> raw;
  raw.reserve(row_bytes * t.height);
  for (size_t s = 0; s < t.strip_offsets.size(); ++s) {
    const size_t rows = std::min<size_t>(t.rows_per_strip, t.height - std::min<size_t>(t.height, s * t.rows_per_strip));
    const size_t want = rows * row_bytes, off = t.strip_offsets[s], cnt = t.strip_counts[s];
    if (off > b.size() || cnt > b.size() - off) fail(file, "strip beyond the end of the file");
    const size_t before = raw.size();
    if (t.compression == 1) raw.insert(raw.end(), b.begin() + off, b.begin() + off + std::min(cnt, want));
    else if (t.compression == 5) lzw_decode(b.data() + off, cnt, raw, want, file);
    else packbits_decode(b.data() + off, cnt, raw, want);
    if (raw.size() - before < want) fail(file, "strip " + std::to_string(s) + " holds too few bytes");
    raw.resize(before + want);
  }
  if (raw.size() < row_bytes * t.height) fail(file, "too few strips");
  std::vector<uint16_t> v((size_t)t.width * t.height);
  for (size_t i = 0; i < v.size(); ++i)
    v[i] = bps == 1 ? raw[i] : t.big_endian ? (uint16_t)(raw[2 * i] << 8 | raw[2 * i + 1]) : (uint16_t)(raw[2 * i] | raw[2 * i + 1] << 8);
  if (t.predictor == 2) {
    const uint32_t mask = bps == 1 ? 0xffu : 0xffffu;
    for (int y = 0; y < t.height; ++y)
      for (int x = 1; x < t.width; ++x) {
        const size_t i = (size_t)y * t.width + x;
        v[i] = (uint16_t)((v[i] + v[i - 1]) & mask);
      }
  }
  width = t.width;
  height = t.height;
  bits = t.bits;
  photometric = t.photometric;
  if (colormap) *colormap = t.colormap;
  return v;
}

bool has(const std::string &s, const char *sub) { return s.find(sub) != std::string::npos; }

}  // namespace

Image_File_Type image_file_type(const std::string &f) {
  if (has(f, ".rawi")) return RAWI;
  if (has(f, ".cine")) return CINE;
  if (has(f, "tif") || has(f, "tiff") || has(f, "TIF") || has(f, "TIFF")) return TIFF;
  if (has(f, "jpg") || has(f, "jpeg") || has(f, "JPG") || has(f, "JPEG")) return JPEG;
  if (has(f, "bmp") || has(f, "BMP")) return BMP;
  if (has(f, "png") || has(f, "PNG")) return PNG;
  if (has(f, ".nc")) return NETCDF;
  if (has(f, ".avi") || has(f, ".mp4")) return VIDEO;
  return NO_SUCH_IMAGE_FILE_TYPE;
}

template <typename T>
void read_rawi_image(const std::string &file, std::vector<T> &values, int &width, int &height) {
  const std::vector<unsigned char> b = slurp(file);
  if (b.size() < 12) fail(file, "truncated .rawi header");
  uint32_t hdr[3];
  std::memcpy(hdr, b.data(), sizeof hdr);
  if (hdr[2] != sizeof(T)) fail(file, "saved with a different value size (" + std::to_string(hdr[2]) + " bytes)");
  const size_t n = (size_t)hdr[0] * hdr[1];
  if (b.size() - 12 < n * sizeof(T)) fail(file, "truncated .rawi data");
  values.resize(n);
  std::memcpy(values.data(), b.data() + 12, n * sizeof(T));
  width = (int)hdr[0];
  height = (int)hdr[1];
}

template <typename T>
void write_rawi_image(const std::string &file, int width, int height, const T *values) {
  if (width < 1 || height < 1 || !values) fail(file, "bad image");
  std::ofstream out(file, std::ios::binary);
  if (!out) fail(file, "cannot create");
  const uint32_t hdr[3] = {(uint32_t)width, (uint32_t)height, (uint32_t)sizeof(T)};
  out.write((const char *)hdr, sizeof hdr);
  out.write((const char *)values, (std::streamsize)((size_t)width * height * sizeof(T)));
  if (!out) fail(file, "write failed");
}

template void read_rawi_image<float>(const std::string &, std::vector<float> &, int &, int &);
template void read_rawi_image<double>(const std::string &, std::vector<double> &, int &, int &);
template void write_rawi_image<float>(const std::string &, int, int, const float *);
template void write_rawi_image<double>(const std::string &, int, int, const double *);

void read_image_dimensions(const std::string &file, int &width, int &height) {
  const Image_File_Type type = image_file_type(file);
  if (type == RAWI) {
    const std::vector<unsigned char> b = slurp(file);
    if (b.size() < 12) fail(file, "truncated .rawi header");
    uint32_t hdr[2];
    std::memcpy(hdr, b.data(), sizeof hdr);
    width = (int)hdr[0];
    height = (int)hdr[1];
  } else if (type == TIFF) {
    const Tiff t = parse_tiff(slurp(file), file);
    width = t.width;
    height = t.height;
  } else {
    fail(file, type == NO_SUCH_IMAGE_FILE_TYPE ? "unrecognized image file type" : "image file type not supported");
  }
}

void read_tiff_samples(const std::string &file, std::vector<uint16_t> &samples, int &width, int &height, int &bits) {
  if (image_file_type(file) != TIFF) fail(file, "not a TIFF file name");
  int photometric = 1;
  samples = tiff_samples(file, width, height, bits, photometric);
}

void read_image(const std::string &file, std::vector<float> &intensities, int &width, int &height, int sub_offset_x,
                int sub_offset_y, int sub_w, int sub_h) {
  const Image_File_Type type = image_file_type(file);
  const bool sub = sub_offset_x || sub_offset_y || sub_w || sub_h;
  if (type == RAWI) {
    if (sub) fail(file, "reading only a portion of an image is not supported for rawi");
    read_rawi_image<float>(file, intensities, width, height);
    return;
  }
  if (type != TIFF) fail(file, type == NO_SUCH_IMAGE_FILE_TYPE ? "unrecognized image file type" : "image file type not supported");
  int w = 0, h = 0, bits = 8, photometric = 1;
  std::vector<uint16_t> cmap;
  const std::vector<uint16_t> v = tiff_samples(file, w, h, bits, photometric, &cmap);
  // palette images: the 16-bit colour map to 8 bits (x >> 8), then to grey with the BT.601
  // weights in 14-bit fixed point (R 4899, G 9617, B 1868, rounded); a grey colour map maps
  // each index to its own grey level exactly
  uint8_t grey[256] = {0};
  if (photometric == 3)
    for (int k = 0; k < 256; ++k) {
      const uint32_t r = cmap[k] >> 8, gg = cmap[256 + k] >> 8, b = cmap[512 + k] >> 8;
      grey[k] = (uint8_t)((r * 4899u + gg * 9617u + b * 1868u + 8192u) >> 14);
    }
  if (sub_offset_x < 0 || sub_offset_y < 0 || sub_w < 0 || sub_h < 0) fail(file, "bad sub-image");
  width = sub_w == 0 ? w : sub_w;
  height = sub_h == 0 ? h : sub_h;
  if (width + sub_offset_x > w || height + sub_offset_y > h) fail(file, "sub-image extends beyond the image");
  intensities.assign((size_t)width * height, 0.0f);
  for (int y = 0; y < height; ++y)
    for (int x = 0; x < width; ++x) {
      uint32_t s = v[(size_t)(y + sub_offset_y) * w + x + sub_offset_x];
      if (bits == 16) s = (s * 255u + 32767u) / 65535u;
      if (photometric == 0) s = 255u - s;
      if (photometric == 3) s = grey[s & 255u];
      intensities[(size_t)y * width + x] = (float)s;
    }
}

}  // namespace dice_hip
