// qkrylov/basis/hubbard_basis.hpp at the path of the reference header of that name; the
// names it declares are in qkrylov/qkrylov.hpp.
#pragma once
#include "../qkrylov.hpp"
