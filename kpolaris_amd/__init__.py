"""MI355X-native KPolaris imaging kernel: Python host layer over the C ABI."""
