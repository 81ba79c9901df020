// Same include path as MTL4; the sparse Krylov path of MTL4 on MI355X.
#pragma once
#include <mtl4g/dropin.hpp>
