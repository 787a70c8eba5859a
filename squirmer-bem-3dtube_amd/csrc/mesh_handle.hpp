// mesh_handle.hpp -- the opaque sqbem_mesh of the C ABI: a host bem::Mesh.
#pragma once
#include "../../include/utils/mesh.hpp"

struct sqbem_mesh {
    bem::Mesh m;
};
