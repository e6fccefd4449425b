// io_common.h -- host-side readers of scene and asset files (no device code): affine
// transforms, meshes, images, and the scene loader built on them.
#pragma once
#include "../../include/sms.h"

#include <cstdint>
#include <string>
#include <vector>

namespace smsio {

// 4x4 row-major affine transform, FP64
struct Mat4 { double a[16]; };
Mat4 mat_identity();
Mat4 mat_mul(const Mat4 &x, const Mat4 &y);
bool mat_inverse(const Mat4 &m, Mat4 &out);
void mat_point(const Mat4 &m, const double p[3], double out[3]);
void mat_vector(const Mat4 &m, const double v[3], double out[3]);

struct MeshData {
    int64_t n_vertices = 0, n_faces = 0;
    std::vector<float> positions, normals, texcoords;   // normals / texcoords empty when absent
    std::vector<uint32_t> faces;
};

struct Image {
    int width = 0, height = 0, channels = 0;
    std::vector<float> pixels;                           // row-major, `channels` floats per pixel
};

bool read_file(const std::string &path, std::vector<unsigned char> &out, std::string &err);
sms_status load_obj(const std::string &path, const Mat4 *to_world, int flags, MeshData &m, std::string &err);
sms_status load_image(const std::string &path, Image &img, std::string &err);

} // namespace smsio
