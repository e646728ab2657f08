// scene_xml.h -- scene files on the host: the Mitsuba-XML subset of mitsuba.rs, Wavefront OBJ
// (obj_loader.rs) and Radiance RGBE images (hdr.rs read side), producing a tpt_scene_desc.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tinypt_gpu.h"

namespace tpt {

// Sets tpt_last_error() (defined with the rest of the C ABI).
int set_last_error(int code, const std::string& msg);

// OBJ triangles, 9 doubles each (obj_loader.rs:25-69, 101-109).
bool load_obj(const std::string& path, std::vector<double>* tris, std::string* err);

// Radiance RGBE image as linear RGB, 3 doubles per pixel, rows top to bottom (hdr.rs:20-62).
bool load_hdr(const std::string& path, int* w, int* h, std::vector<double>* rgb, std::string* err);

}  // namespace tpt
