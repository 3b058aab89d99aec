"""MI355X-native thick-restart Lanczos (nu-TRLan trlan/ztrlan drop-in)."""
