/* hd/parse_command_line.h -- replaces Pipeline/hd/parse_command_line.h:16-17 */
#pragma once
#include "hd/params.h"
#ifdef __cplusplus
extern "C" {
#endif
/* replaces hd_parse_command_line (Pipeline/parse_command_line.C:23-158) */
int  hd_parse_command_line(int argc, char* argv[], struct hd_params* params);
/* replaces hd_print_usage (Pipeline/parse_command_line.C:160-188) */
void hd_print_usage(void);
#ifdef __cplusplus
}
#endif
