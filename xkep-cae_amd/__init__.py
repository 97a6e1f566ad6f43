"""MI355X-native per-Newton-iteration assembly path for strand contact (xkep-cae drop-in)."""
