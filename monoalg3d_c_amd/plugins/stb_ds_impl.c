/*
 * The element lists the assembly plugin writes are stb_ds arrays that the
 * core later reads and frees: the plugin library carries its own copy of the
 * array functions, as each of the reference's plugins does (they link the
 * core's objects, where utils/file_utils.c compiles them).
 */
#define STB_DS_IMPLEMENTATION
#include "3dparty/stb_ds.h"
