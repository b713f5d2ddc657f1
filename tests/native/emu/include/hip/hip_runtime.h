#pragma once
#include "../../emu_runtime.h"
