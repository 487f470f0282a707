/*
 * Forwarding header for code written against the reference's private header
 * (include/dplr_internal.h of dplr): the node types and the private
 * lifecycle routines it declares -- dplrinid, dplrinilr, dplrinih and
 * dplrresizelr (reference dplr_internal.h:32-87 and :99-115) -- are
 * declared, with the same layout and prototypes, in dplr.h of this library,
 * so a translation unit that includes this header builds unchanged.
 */
#ifndef DPLR_INTERNAL_FORWARD_H
#define DPLR_INTERNAL_FORWARD_H

#include "dplr.h"

#endif
