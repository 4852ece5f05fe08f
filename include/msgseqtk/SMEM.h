/* SMEM.h -- the SMEM class of the drop-in face (msgseqtk_api.h) */
#include "msgseqtk_api.h"
