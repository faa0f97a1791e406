// Binary STL reading and writing (src/io/stl.rs), host only.
#pragma once
#include <cstddef>
#include <string>
#include <vector>

namespace bs {

// StlReader::read_from_buffer (io/stl.rs:64-92): 80-byte header, u32 face
// count, per face a normal (skipped), three vertices (f32 little endian) and a
// u16 attribute.  Appends 9 floats per face to `xyz`; returns an empty string
// on success, else the reason (a short file is an error, as read_exact's).
std::string stl_read(const char* path, std::vector<float>& xyz);

// StlWriter::write_to_buffer (io/stl.rs:150-199): zero header, u32 count, per
// face the unit normal of (b - a) x (c - a) (zeros when its squared norm is 0,
// triangle3.rs:261-269), the three vertices and a zero attribute.
std::string stl_write(const char* path, const float* xyz, size_t n_tris);

}  // namespace bs
