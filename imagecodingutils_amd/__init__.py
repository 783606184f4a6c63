"""MI355X-native motion estimation and H.264 4x4 transform/quant for ImageCodingUtils."""
