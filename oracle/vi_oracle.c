/*
 * vi_oracle.c - CPU restatement of the VI scanout filters (TEST
 * INFRASTRUCTURE ONLY, see rdp_oracle.h).
 */
#include "rdp_oracle.h"

int rdpo_scanout(rdpo *o, uint32_t *out, unsigned *width, unsigned *height)
{
	(void)o;
	(void)out;
	*width = 0;
	*height = 0;
	return -1;
}
