#pragma once
#include "lammps_stub.h"
