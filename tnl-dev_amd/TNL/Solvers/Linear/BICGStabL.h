#pragma once
#include "LinearSolver.h"
