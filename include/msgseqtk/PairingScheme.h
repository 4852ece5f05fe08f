/* PairingScheme.h -- the PairingScheme and AlignmentPE classes of the drop-in face (msgseqtk_api.h) */
#include "msgseqtk_api.h"
