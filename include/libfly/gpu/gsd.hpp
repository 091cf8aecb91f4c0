// fly::io::BinaryFile, GSD trajectory files (the reference's libfly/io/gsd.hpp): defined with the C++
// drop-in, whose minimiser and dimer write frames through it (Options::fout).
#pragma once

#include "facade.hpp"
