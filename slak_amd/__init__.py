"""slak_amd: MI355X-native Catmull-Clark subdivision (LAK / SLAK) behind a C-ABI."""
